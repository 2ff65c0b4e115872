// Gaussian (mean, sigma) convolved with exp(-alpha t): the exponentially
// modified Gaussian (include/goofit/PDFs/basic/ExpGausPdf.h of the reference);
// normalised numerically on the observable grid.
#pragma once

#include <goofit/PDFs/GooPdf.h>

namespace GooFit {

class ExpGausPdf : public GooPdf {
  public:
    ExpGausPdf(std::string n, Observable _x, Variable m, Variable s, Variable t);
};

}  // namespace GooFit
