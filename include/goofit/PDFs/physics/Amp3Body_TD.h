// Time-dependent three-body Dalitz-plot PDF (include/goofit/PDFs/physics/Amp3Body_TD.h
// of the reference): a D0 -> abc decay that proceeds directly or through D0-D0bar
// mixing,
//     P(m12, m13, t, sigma_t) = e^{-t/tau} [ (|A|^2 + |B|^2) cosh(y t/tau) + (|A|^2 - |B|^2) cos(x t/tau)
//                                          - 2 Re(A B*) sinh(y t/tau) - 2 Im(A B*) sin(x t/tau) ]
// convolved with a decay-time resolution and multiplied by an efficiency, where
// A = sum_i c_i B_i(m12, m13) and B = sum_i c_i B_i(m13, m12).  The lineshapes of
// the data events are cached on the GPU (both orders, recomputed only when a
// resonance's mass or width changes) and the Dalitz integrals of a_i a_j*,
// a_i b_j*, b_i b_j* over the m12 x m13 grid come from one FP64-MFMA Gram kernel
// of the stacked waves.
#pragma once

#include <complex>
#include <vector>

#include <goofit/PDFs/physics/Amp3BodyBase.h>
#include <goofit/PDFs/physics/DalitzPlotHelpers.h>
#include <goofit/PDFs/physics/MixingTimeResolution.h>
#include <goofit/PDFs/physics/resonances/Resonance.h>

namespace GooFit {

class Amp3Body_TD : public Amp3BodyBase {
  public:
    /// Observables dtime, sigma_t, m12, m13, the event number, then optionally the
    /// mistag probability and the charm tag (+1 D0, -1 D0bar, other: no tag).
    /// Without a mistag observable the mistag terms are off; the reference
    /// requires both pointers.
    Amp3Body_TD(std::string n, Observable _dtime, Observable _sigmat, Observable m12, Observable m13,
                EventNumber eventNumber, DecayInfo3t decay, MixingTimeResolution *r, GooPdf *eff,
                Observable *mistag = nullptr, Observable *charmtag = nullptr);
    /// The reference's variant with one resolution per bin of the D0 mass; its
    /// device function never selects among them (Amp3Body_TD.cu:176-205 are
    /// commented out), so this constructor is refused with an error.
    Amp3Body_TD(std::string n, Observable _dtime, Observable _sigmat, Observable m12, Observable m13,
                EventNumber eventNumber, DecayInfo3t decay, std::vector<MixingTimeResolution *> &r, GooPdf *eff,
                Observable md0, Observable *mistag = nullptr, Observable *charmtag = nullptr);

    fptype normalize() override;
    /// Allocate the wave caches for dataSize events starting at row offset
    /// (call after setData).  evtSize counts the event's columns (at least 5).
    void setDataSize(unsigned int dataSize, unsigned int evtSize = 5, unsigned int offset = 0);
    void setD0Fraction(fptype d0fraction);
    fptype getD0Fraction();
    void setForceIntegrals(bool f = true) { forceRedoIntegrals = f; }
    static void resetCacheCounter() { cacheCount = 0; }

    DecayInfo3t &getDecayInfo() { return decayInfo; }
    /// Host copy of the cached (A, B) waves of resonance i.
    std::vector<WaveHolder_s> getCachedWaveNoCopy(size_t i) const;
    /// Integral contribution of each resonance and interference term: prints the
    /// fractions and returns Re(c_i c_j* Iaa_ij) (Amp3Body_TD.cu:754-910).
    std::vector<std::vector<fptype>> getFractions();

    /// Dalitz integrals of the last normalize(): sum eff a_i a_j*, a_i b_j*, b_i b_j* (n x n).
    const std::vector<std::complex<fptype>> &getIntegralsAA() const { return iaa; }
    const std::vector<std::complex<fptype>> &getIntegralsAB() const { return iab; }
    const std::vector<std::complex<fptype>> &getIntegralsBB() const { return ibb; }

    void populateArrays() override;

  private:
    void computeIntegrals(uint64_t mask);

    DecayInfo3t decayInfo;
    Observable _m12;
    Observable _m13;
    MixingTimeResolution *resolution{nullptr};
    bool hasMistag{false};
    Observable _mistag;

    std::vector<std::complex<fptype>> iaa, iab, ibb;
    std::vector<bool> redoIntegral;
    mutable bool forceRedoIntegrals{true};

    unsigned int resolutionFunction{0};
    unsigned int efficiencyFunction{0};
    int totalEventSize{5};
    int eventOffset{0};
    int cacheToUse{0};
    static int cacheCount;
    bool cacheReady{false};
    fptype _D0Fraction{0.5};
};

}  // namespace GooFit
