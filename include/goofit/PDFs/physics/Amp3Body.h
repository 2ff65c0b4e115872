// Time-independent three-body Dalitz-plot PDF (include/goofit/PDFs/physics/Amp3Body.h
// of the reference):
//     P(m12^2, m13^2) = |sum_i c_i B_i(m12^2, m13^2)|^2 * eff(m12^2, m13^2)
// with complex couplings c_i and resonance lineshapes B_i.  The lineshapes of the
// data events are cached on the GPU and recomputed only when a resonance's mass
// or width changes; the normalisation integrals I_ij = sum eff B_i conj(B_j)
// over the m12 x m13 bin grid come from one FP64-MFMA Gram kernel.
#pragma once

#include <complex>
#include <vector>

#include <goofit/PDFs/physics/Amp3BodyBase.h>
#include <goofit/PDFs/physics/DalitzPlotHelpers.h>
#include <goofit/PDFs/physics/resonances/Resonance.h>

namespace GooFit {

class DalitzProposal;

class Amp3Body : public Amp3BodyBase {
  public:
    Amp3Body(std::string n, Observable m12, Observable m13, EventNumber eventNumber, DecayInfo3 decay,
             GooPdf *eff = nullptr);

    fptype normalize() override;
    /// The coupling quadratic form over the integrals of the last normalize() as a device
    /// op (the resonance masses and widths must stay fixed while the op is in use).
    bool batchNormOps(std::vector<gf_norm_op> &ops, std::vector<double> &table) const override;
    /// Allocate the wave cache for dataSize events starting at row offset
    /// (call after setData).
    void setDataSize(unsigned int dataSize, unsigned int evtSize = 3, unsigned int offset = 0);
    void setForceIntegrals(bool f = true) { forceRedoIntegrals = f; }
    /// The next normalize() recomputes the wave cache of every resonance from the
    /// current events but keeps the integrals of the last one, unless a resonance
    /// parameter moved: for fitting one model to many data sets in turn (call after
    /// setDataSize).  No counterpart in the reference, which recomputes both.
    void refreshCacheOnly() {
        forceRedoIntegrals = true;
        cacheOnly          = true;
    }

    void populateArrays() override;

    /// Host copy of the cached lineshape of resonance i.
    const std::vector<std::complex<fptype>> getCachedWave(size_t i) const;
    /// Sum of the cached lineshape of resonance i over the events.
    std::complex<fptype> sumCachedWave(size_t i) const;

    DecayInfo3 &getDecayInfo() { return decayInfo; }
    static void resetCacheCounter() { cacheCount = 0; }

    /// Fit fractions from the data-event cache (cache must be filled).
    std::vector<std::vector<fptype>> fit_fractions();

    /// Integral of |sum_i c_i W_i|^2 eff over the plot by the quadrature (m12[k], m13[k],
    /// w[k]) at the current couplings: the integrals I_ij as one Gram over the points,
    /// then the quadratic form of normalize() (no bin-area factor).  For judging a
    /// normalisation method against a converged rule; no counterpart in the reference.
    fptype integralAtPoints(const std::vector<fptype> &m12, const std::vector<fptype> &m13,
                            const std::vector<fptype> &w);

    /// Normalisation integrals I_ij of the last normalize() (n x n).
    const std::vector<std::complex<fptype>> &getIntegrals() const { return integrals; }
    /// Function index of the efficiency (the last function of this subtree).
    unsigned int getEfficiencyIndex() const { return efficiencyFunction; }

    /// Normalise by n Monte-Carlo points of the (m12, m13) box instead of the midpoint
    /// grid: uniform (proposal == nullptr) or drawn from a learned proposal with
    /// importance weights.  The integrals are then sum_p w_p eff B_i conj(B_j) (no bin
    /// area factor).  No counterpart in the reference.
    void setNormalisationPoints(int64_t n, unsigned long long seed, DalitzProposal *proposal = nullptr);
    /// Normalise by a deterministic quadrature rule over the plot (points m12[k], m13[k]
    /// with weights w[k], e.g. a product Gauss-Legendre rule fitted to the kinematic
    /// boundary): the integrals are sum_k w_k eff B_i conj(B_j) (no bin-area factor).
    /// Unlike the midpoint grid, no cell straddles the boundary.  No counterpart in the
    /// reference.
    void setNormalisationRule(const std::vector<fptype> &m12, const std::vector<fptype> &m13,
                              const std::vector<fptype> &w);
    /// Back to the reference's midpoint grid (the default).
    void setNormalisationGrid();
    int64_t getNormalisationPointCount() const { return normPoints; }
    /// Relative variance per point of the normalisation estimate at the current
    /// couplings, from `batches` disjoint batches of the points (Monte-Carlo mode).
    fptype normalisationRelativeVariance(int batches = 16);

    /// Append n events of this PDF's density at the current parameters to `out`
    /// (observables m12, m13 and the event number, numbered from out's current
    /// size), generated on the GPU by accept-reject with a seeded counter-based
    /// stream.  fmax <= 0 estimates the envelope.  Returns the candidates tried.
    /// Takes the place of the reference's MCBooster-based GenerateSig
    /// (src/PDFs/physics/Amp3Body.cu:430-547).
    int64_t generateEvents(UnbinnedDataSet &out, size_t n, unsigned long long seed, double fmax = 0);

  protected:
    DecayInfo3 decayInfo;
    Observable _m12;
    Observable _m13;
    EventNumber _eventNumber;

    std::vector<std::complex<fptype>> integrals;
    std::vector<bool> redoIntegral;
    mutable bool forceRedoIntegrals{true};
    int totalEventSize{3};
    int eventOffset{0};
    int cacheToUse{0};
    static int cacheCount;
    unsigned int efficiencyFunction{0};
    bool cacheReady{false};
    int64_t normPoints{0};  // 0: midpoint grid
    bool cacheOnly{false};     // refreshCacheOnly: the next normalize() keeps the integrals
    bool haveIntegrals{false};  // the integrals have been computed at least once
};

}  // namespace GooFit
