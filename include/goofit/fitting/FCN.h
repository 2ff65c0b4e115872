// Minuit2 function wrapping GooFit's calculateNLL (include/goofit/fitting/FCN.h of the reference).
#pragma once

#include <vector>

#include <Minuit2/FCNBase.h>
#include <goofit/fitting/Params.h>

namespace GooFit {

class FCN : public Minuit2::FCNBase {
  protected:
    Params *params_;

  public:
    FCN(Params &params);
    /// NLL at the Minuit parameter vector pars.
    double operator()(const std::vector<double> &pars) const override;
    /// NLL at the current parameter values.
    double operator()() const;
    /// 0.5 for -log L.
    double Up() const override { return 0.5; }
    Params *GetParams();
};

}  // namespace GooFit
