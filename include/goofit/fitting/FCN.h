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
    /// NLL at several Minuit parameter vectors, in order: each point's parameters
    /// are set and normalised as operator() does, then all metric sums come from
    /// one gf_reduce_metric_multi request (one sweep per 16 points for a lone
    /// Amp3Body).  Points whose normalisation recomputes a wave cache are
    /// evaluated one by one.  The values equal operator()'s.
    std::vector<double> EvaluateMany(const std::vector<std::vector<double>> &xs) const override;
    /// Serve multi-point requests together (default) or one by one.
    void setBatched(bool on) { batched_ = on; }
    bool isBatched() const { return batched_; }
    /// 0.5 for -log L.
    double Up() const override { return 0.5; }
    Params *GetParams();

  private:
    bool batched_{true};
};

}  // namespace GooFit
