// Polynomials in one or several observables
// (include/goofit/PDFs/basic/PolynomialPdf.h of the reference).
#pragma once

#include <memory>

#include <goofit/PDFs/GooPdf.h>

namespace GooFit {

class PolynomialPdf : public GooPdf {
  public:
    PolynomialPdf(std::string n, Observable _x, std::vector<Variable> weights, unsigned int lowestDegree = 0);
    PolynomialPdf(std::string n, Observable _x, std::vector<Variable> weights, Variable x0,
                  unsigned int lowestDegree = 0);
    PolynomialPdf(std::string n, std::vector<Observable> obses, std::vector<Variable> coeffs,
                  std::vector<Variable> offsets, unsigned int maxDegree = 0);
    fptype integrate(fptype lo, fptype hi) const override;
    fptype getCoefficient(int coef);

  private:
    std::unique_ptr<Variable> center;
};

}  // namespace GooFit
