// Weighted sum of PDFs, extended or not (include/goofit/PDFs/combine/AddPdf.h of the reference).
#pragma once

#include <goofit/PDFs/CombinePdf.h>

namespace GooFit {

class AddPdf : public CombinePdf {
  public:
    AddPdf(std::string n, std::vector<Variable> weights, std::vector<PdfBase *> comps);
    AddPdf(std::string n, Variable frac1, PdfBase *func1, PdfBase *func2);
    fptype normalize() override;
    double calculateNLL() override;
    double nllExtraTerm() const override;
    double nllExtraTermFor(fptype nevents) const override;
    bool hasAnalyticIntegral() const override { return false; }
    bool batchNormOps(std::vector<gf_norm_op> &ops, std::vector<double> &table) const override;
    std::vector<unsigned int> extendedYieldSlots() const override;

  private:
    bool extended;
};

}  // namespace GooFit
