// Decay-time resolution models of the time-dependent Dalitz PDF
// (include/goofit/PDFs/physics/MixingTimeResolution.h of the reference).  A
// resolution is a component of Amp3Body_TD: the engine evaluates it inside the
// Amp3Body_TD node with the node's rate terms, and normalization() gives its
// decay-time integral for the host assembly of the PDF's normalisation.
#pragma once

#include <goofit/PDFs/GooPdf.h>

namespace GooFit {

class MixingTimeResolution : public GooPdf {
  public:
    template <class... Args>
    MixingTimeResolution(std::string pdf_name, Args &&...args)
        : GooPdf(pdf_name, "mixing_resolution", std::forward<Args>(args)...) {}

    ~MixingTimeResolution() override;

    /// Integral over the decay time of the rate with Dalitz integrals di1 = |A|^2,
    /// di2 = |B|^2, di3 + i di4 = A B*.
    virtual fptype
    normalization(fptype di1, fptype di2, fptype di3, fptype di4, fptype tau, fptype xmixing, fptype ymixing) const
        = 0;

    /// The gf_func code of the resolution's device function (positive once registered).
    int getDeviceFunction() const { return function_code_ > 0 ? function_code_ : -1; }
    int getCalcTauIdx() const { return resCalcTauFcnIdx; }
    void setCalcTauIdx(int idx) { resCalcTauFcnIdx = idx; }

  private:
    int resCalcTauFcnIdx{-1};
};

}  // namespace GooFit
