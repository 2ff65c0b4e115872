// A resonance lineshape B_i of a three-body amplitude
// (include/goofit/PDFs/physics/resonances/Resonance.h of the reference).  Its
// complex coefficient (amp_real, amp_imag) is a parameter of the parent Amp3Body.
#pragma once

#include <goofit/PDFs/physics/AmpComponent.h>
#include <goofit/PDFs/physics/DalitzPlotHelpers.h>

namespace GooFit {

class ResonancePdf : public AmpComponent {
    friend class Amp3Body;
    friend class Amp3Body_TD;

  public:
    ~ResonancePdf() override = default;
    virtual void recalculateCache() const {}
    Variable get_amp_real() const { return amp_real; }
    Variable get_amp_img() const { return amp_imag; }

  protected:
    ResonancePdf(std::string pdf_name, std::string name, Variable ar, Variable ai)
        : AmpComponent("Resonances::" + pdf_name, name)
        , amp_real(ar)
        , amp_imag(ai) {}

    Variable amp_real;
    Variable amp_imag;
};

}  // namespace GooFit
