// Learned importance-sampling proposal for the normalisation integrals of an
// Amp3Body (no counterpart in the reference, which integrates on a fixed
// midpoint grid, src/PDFs/physics/detail/SpecialResonanceIntegrator.cu:52-121).
//
// A 2 -> 128 -> 128 -> 1 SiLU network f over the (m12, m13) box is trained on the
// GPU (BF16 MFMA, FP32 accumulation) by maximum likelihood on events the model
// generated; exp(f) tabulated on a grid and mixed with the uniform density is the
// proposal q.  Normalisation points drawn from q carry weights 1 / (N q), so the
// integrals stay unbiased whatever the network learned; uniform phase-space Monte
// Carlo is the fallback (Amp3Body::setNormalisationPoints without a proposal).
#pragma once

#include <utility>
#include <vector>

#include <goofit/UnbinnedDataSet.h>
#include <goofit/Variable.h>
#include <goofit_hip.h>

namespace GooFit {

class Amp3Body;

class DalitzProposal {
  public:
    /// A network with seeded initial weights over the box of m12 x m13 (their limits).
    DalitzProposal(Observable m12, Observable m13, unsigned long long seed = 1, int batch_cap = 65536);
    ~DalitzProposal();
    DalitzProposal(const DalitzProposal &)            = delete;
    DalitzProposal &operator=(const DalitzProposal &) = delete;

    /// Maximum-likelihood training on the (m12, m13) columns of events; returns the
    /// loss of the first and of the last step.
    std::pair<float, float> train(const UnbinnedDataSet &events, int steps = 2000, int batch = 8192,
                                  float lr = 3e-3f, unsigned long long seed = 1);
    /// The same training on nevents events of the model's density at its current
    /// parameters, generated on the GPU and kept there (no host copy of the events).
    std::pair<float, float> trainOnModel(Amp3Body &model, int64_t nevents, unsigned long long gen_seed,
                                         int steps = 2000, int batch = 8192, float lr = 3e-3f,
                                         unsigned long long seed = 1);
    /// Tabulate the proposal on grid x grid cells with uniform mixture weight alpha.
    void build(int grid = 512, double alpha = 0.1);
    bool isBuilt() const { return built_; }

    /// Network output f at points (x, y) in [-1, 1]^2 (interleaved), u = (x + 1) / 2 of the box.
    std::vector<float> evaluate(const std::vector<float> &xy) const;
    std::vector<float> getParameters() const;
    void setParameters(const std::vector<float> &p);

    /// Retraining trigger: the relative weight variance per point measured when the
    /// proposal was built for the current model (Amp3Body::normalisationRelativeVariance),
    /// and the test relvar > factor * reference.
    void setReferenceVariance(double v) { ref_var_ = v; }
    double getReferenceVariance() const { return ref_var_; }
    bool needsRetraining(double relvar, double factor = 4.0) const { return ref_var_ > 0 && relvar > factor * ref_var_; }

    /// Automatic retraining by Amp3Body::setNormalisationPoints: each time points are
    /// drawn from this proposal, their relative weight variance is measured at the
    /// current couplings; the first measurement becomes the reference, and a later one
    /// above factor x the reference retrains the network on `events` events that the
    /// model generates at its current parameters (train(steps, batch, lr), build(grid,
    /// alpha)), draws the points again and takes their variance as the new reference.
    struct Retraining {
        int64_t events{2000000};
        int steps{500};
        int batch{8192};
        float lr{6e-3f};
        int grid{512};
        double alpha{0.1};
        double factor{4.0};
    };
    void setRetraining(const Retraining &r);
    void disableRetraining() { retrain_on_ = false; }
    bool retrainingEnabled() const { return retrain_on_; }
    const Retraining &retraining() const { return retrain_; }
    /// Retrains done by the trigger so far.
    int retrainCount() const { return retrains_; }
    /// Train and build again on events of the model (one retraining of the trigger).
    void retrain(Amp3Body &model, unsigned long long seed);

    const double *box() const { return box_; }
    gf_mlp *handle() const { return mlp_; }

  private:
    Observable m12_, m13_;
    double box_[4];
    gf_mlp *mlp_{nullptr};
    bool built_{false};
    double ref_var_{0.0};
    bool retrain_on_{false};
    Retraining retrain_;
    int retrains_{0};
};

}  // namespace GooFit
