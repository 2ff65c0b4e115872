// Perfect decay-time measurement (include/goofit/PDFs/physics/TruthResolution.h
// of the reference): the rate e^{-t/tau} [ (|A|^2+|B|^2) cosh(y t/tau) +
// (|A|^2-|B|^2) cos(x t/tau) - 2 Re(AB*) sinh(y t/tau) - 2 Im(AB*) sin(x t/tau) ].
#pragma once

#include <goofit/PDFs/physics/MixingTimeResolution.h>

namespace GooFit {

class TruthResolution : public MixingTimeResolution {
  public:
    TruthResolution();
    ~TruthResolution() override;

    fptype normalization(fptype di1, fptype di2, fptype di3, fptype di4, fptype tau, fptype xmixing,
                         fptype ymixing) const override;
};

}  // namespace GooFit
