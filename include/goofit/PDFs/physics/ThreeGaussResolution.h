// Three-Gaussian decay-time resolution (include/goofit/PDFs/physics/ThreeGaussResolution.h
// of the reference): core, tail and outlier Gaussians (fractions cf, tf and
// 1 - cf - tf) with biases and scale factors in units of the per-event error
// sigma_t, a selection bias, and common bias offsets and width scales (the
// first pair for the core, the second for tail and outlier).  The rate is
// convolved analytically; cos and sin are expanded to second and third order.
#pragma once

#include <goofit/PDFs/physics/MixingTimeResolution.h>

namespace GooFit {

class ThreeGaussResolution : public MixingTimeResolution {
  public:
    ThreeGaussResolution(Variable cf, Variable tf, Variable cb, Variable cs, Variable tb, Variable ts, Variable ob,
                         Variable os, Variable sb, Variable common_bias_offset, Variable common_width_scale,
                         Variable common_bias_offset2, Variable common_width_scale2);
    ~ThreeGaussResolution() override;

    fptype normalization(fptype di1, fptype di2, fptype di3, fptype di4, fptype tau, fptype xmixing,
                         fptype ymixing) const override;

  private:
    Variable selectionBias;
};

}  // namespace GooFit
