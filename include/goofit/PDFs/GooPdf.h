// GooPdf: a PDF evaluated by the GPU engine (include/goofit/PDFs/GooPdf.h of the
// reference).  The reductions that the reference runs as thrust algorithms over
// MetricTaker are single calls into the C ABI of include/goofit_hip.h.
#pragma once

#include <memory>
#include <vector>

#include <goofit/FitControl.h>
#include <goofit/PdfBase.h>
#include <goofit/UnbinnedDataSet.h>
#include <goofit_hip.h>

namespace GooFit {

class GooPdf : public PdfBase {
  protected:
    /// Number of event columns (negative for binned data: -(observables + 2)).
    int get_event_size() const;
    /// Sum over the events of metric(pdf(event), norm) (does not normalise).
    double reduce_with_metric() const;
    /// Metric of every event (does not normalise).
    void evaluate_with_metric(std::vector<fptype> &results) const;
    std::vector<fptype> evaluate_with_metric() const;
    /// gf_metric code of the current fit control.
    int metric_code() const;

  public:
    using PdfBase::PdfBase;
    ~GooPdf() override = default;

    /// Number of cells of the normalisation grid.
    int get_bin_grid_size() const;
    /// Volume of the normalisation box.
    fptype get_bin_grid_volume() const;
    /// Sum of the raw PDF over the normalisation grid (does not normalise).
    double reduce_with_bins() const;

    double calculateNLL() override;
    /// The metric code (gf_metric) this PDF's fit control selects.
    int metricCode() const { return metric_code(); }
    /// Term calculateNLL adds to the metric sum (the extended-likelihood term of an
    /// extended AddPdf), at the current parameters.
    virtual double nllExtraTerm() const { return 0.0; }
    /// The same term for a data set of nevents events (a toy's share of the rows).
    virtual double nllExtraTermFor(fptype) const { return 0.0; }
    /// The normalisations of this subtree that floating parameters move, as device ops
    /// for batched points (gf_batch_set_norm_ops): appends them (integrals they read go
    /// to table) and returns true, or returns false when one of them has no such form (a
    /// numerical integral over floating parameters).  Call after normalize(): the ops
    /// reproduce the normalisations normalize() computes from the current arrays.  No
    /// counterpart in the reference, which normalises on the host at every call.
    virtual bool batchNormOps(std::vector<gf_norm_op> &ops, std::vector<double> &table) const;
    /// Parameter-array slots whose values sum to the expected event count of the
    /// extended-likelihood term of nllExtraTermFor (empty: no such term).
    virtual std::vector<unsigned int> extendedYieldSlots() const { return {}; }
    /// The normalisation of this node is positive (calculateNLL aborts otherwise).
    bool normalisationPositive() const;
    std::vector<fptype> evaluateAtPoints(Observable var);
    fptype normalize() override;
    fptype normalise() { return normalize(); }

    virtual fptype integrate(fptype lo, fptype hi) const;
    bool hasAnalyticIntegral() const override;
    fptype getValue(EvalFunc evalfunc = EvalFunc::Eval);

    /// Metric of every event of the current data (normalised first), e.g. the
    /// probability of each event for EvalFunc::Prob.  The reference keeps this
    /// as the protected evaluate_with_metric(); exposed for event generation.
    std::vector<fptype> evaluateAtData(EvalFunc evalfunc = EvalFunc::Prob);
    /// Probabilities of the PDF and of each component at the data points.
    std::vector<std::vector<fptype>> getCompProbsAtDataPoints();
    /// Dataset of the bin centres of the PDF's observables.
    UnbinnedDataSet makeGrid();

    void initialize();
    void scan(Observable var, std::vector<fptype> &values);
    void setFitControl(std::shared_ptr<FitControl> fc) override;
    virtual void setMetrics() {}
    void setParameterConstantness(bool constant = true);
    void setIndices() override;
};

}  // namespace GooFit
