// Product of PDFs (include/goofit/PDFs/combine/ProdPdf.h of the reference).
#pragma once

#include <goofit/PDFs/CombinePdf.h>

namespace GooFit {

class ProdPdf : public CombinePdf {
  public:
    ProdPdf(std::string n, std::vector<PdfBase *> comps);
    fptype normalize() override;
    bool hasAnalyticIntegral() const override { return false; }
    bool batchNormOps(std::vector<gf_norm_op> &ops, std::vector<double> &table) const override;

  private:
    bool varOverlaps;
};

}  // namespace GooFit
