// exp(-(x-m)^2 / (2 s^2)), analytic integral s*sqrt(2 pi) over all x
// (include/goofit/PDFs/basic/GaussianPdf.h of the reference).
#pragma once

#include <goofit/PDFs/GooPdf.h>

namespace GooFit {

class GaussianPdf : public GooPdf {
  public:
    GaussianPdf(std::string n, Observable _x, Variable m, Variable s);
    fptype integrate(fptype lo, fptype hi) const override;
    bool hasAnalyticIntegral() const override { return true; }
    bool batchNormOps(std::vector<gf_norm_op> &ops, std::vector<double> &table) const override;
};

}  // namespace GooFit
