// exp(alpha (x - x0)), or exp of a polynomial in (x - x0)
// (include/goofit/PDFs/basic/ExpPdf.h of the reference).
#pragma once

#include <goofit/PDFs/GooPdf.h>

namespace GooFit {

class ExpPdf : public GooPdf {
  public:
    ExpPdf(std::string n, Observable _x, Variable alpha);
    ExpPdf(std::string n, Observable _x, std::vector<Variable> &weights);
    ExpPdf(std::string n, Observable _x, Variable alpha, Variable offset);
    ExpPdf(std::string n, Observable _x, std::vector<Variable> &weights, Variable offset);
    fptype integrate(fptype lo, fptype hi) const override;
    bool hasAnalyticIntegral() const override;
    bool batchNormOps(std::vector<gf_norm_op> &ops, std::vector<double> &table) const override;
};

}  // namespace GooFit
