// Many toy fits of one PDF model at once (no counterpart in the reference,
// which fits one data set per FitManager).
//
// Toy t owns rows [first[t], first[t] + count[t]) of one data set.  One Minuit2
// MIGRAD per toy runs on a pool of host threads; every FCN call, and every group
// of independent points (one derivative round, up to 64 points), posts the toy's
// parameter vectors to a batcher, which turns the pending requests of all toys
// into one GPU launch sequence (staging copy, per-toy event sweep over all of a
// toy's points, per-toy finishing reduction) and hands each toy its NLLs; when
// preparing the requests costs more than the GPU work, a round goes out in two
// stages so that one is evaluated while the other is prepared.  With
// useMultiPoint(false) every point is its own request, served by a captured
// hipGraph with one parameter set per toy.  The normalisations of the points are
// computed inside the batched launch when the model allows it
// (GooPdf::batchNormOps: analytic integrals, an Amp3Body's quadratic form over its
// cached integral matrix), otherwise on the host per point; either way the
// model's non-analytic parts must not depend on the floating parameters: fit()
// refuses a floating resonance mass or width.
#pragma once

#include <cstdint>
#include <vector>

#include <goofit/GlobalCudaDefines.h>
#include <goofit_hip.h>

namespace GooFit {

class GooPdf;
class UnbinnedDataSet;

struct BatchFitStats {
    int64_t launches   = 0;  // batched GPU reductions
    int64_t fcn_calls  = 0;  // FCN points served
    double seconds     = 0;  // wall time of fit()
    double gpu_wait_s  = 0;  // time spent in the batched reductions
    double host_prep_s = 0;  // time the batcher spends staging each served toy's parameters (and norms)
    double gather_s    = 0;  // time the batcher waits for the toys' next FCN requests
    int chunk          = 0;  // events per workgroup chosen by the autotuner
    bool device_norms  = false;  // the points' normalisations were computed in the launches
    bool two_stages    = false;  // rounds went out in two pipelined stages
};

class BatchFitter {
  public:
    /// pdf must already hold `data` (setData; for an Amp3Body also setDataSize).
    BatchFitter(GooPdf *pdf, const std::vector<int64_t> &first, const std::vector<int64_t> &count);
    ~BatchFitter();
    BatchFitter(const BatchFitter &)            = delete;
    BatchFitter &operator=(const BatchFitter &) = delete;

    /// Fit every toy with MIGRAD from the Variables' current values using `threads`
    /// concurrent Minuit instances (strategy, maxfcn 0 = Minuit's default, tolerance).
    void fit(int threads = 256, unsigned int strategy = 1, double tolerance = 0.1, bool autotune = true);

    /// NLL of every toy at the Variables' current values (one batched launch).
    std::vector<double> evaluate();

    size_t size() const { return first_.size(); }
    /// Per toy: fitted values / errors of the Minuit parameters (Minuit order), validity, calls.
    const std::vector<std::vector<double>> &values() const { return values_; }
    const std::vector<std::vector<double>> &errors() const { return errors_; }
    const std::vector<int> &valid() const { return valid_; }
    const std::vector<int> &nfcn() const { return nfcn_; }
    const std::vector<double> &fval() const { return fval_; }
    const BatchFitStats &stats() const { return stats_; }

    void useGraph(bool on);
    /// Several points per toy and launch (default) or one.
    void useMultiPoint(bool on) { multi_ = on; }
    bool multiPoint() const { return multi_; }
    /// With several points per launch: normalise the points on the device when the
    /// model allows it (default), or always on the host.
    void useDeviceNormalisation(bool on) { device_norms_ = on; }
    /// Rounds of several points per toy in one stage (1), two pipelined stages (2), or
    /// decided from the first rounds' preparation and GPU times (0, default).
    void setStages(int n) { stages_ = (n == 1 || n == 2) ? n : 0; }

  private:
    GooPdf *pdf_;
    std::vector<int64_t> first_, count_;
    gf_batch *batch_{nullptr};
    std::vector<std::vector<double>> values_, errors_;
    std::vector<int> valid_, nfcn_;
    std::vector<double> fval_;
    BatchFitStats stats_;
    bool multi_{true};
    bool device_norms_{true};
    int stages_{0};
};

}  // namespace GooFit
