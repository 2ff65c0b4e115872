// Three-body kinematics shared by the Dalitz PDFs
// (include/goofit/PDFs/physics/DalitzPlotHelpers.h of the reference).
#pragma once

#include <vector>

#include <goofit/GlobalCudaDefines.h>
#include <goofit/Variable.h>

namespace GooFit {

class ResonancePdf;

enum DaughterPair { PAIR_12 = 0, PAIR_13, PAIR_23 };

/// Kinematic boundary of the (m12^2, m13^2) plane.
bool inDalitz(const fptype &m12, const fptype &m13, const fptype &bigM, const fptype &dm1, const fptype &dm2,
              const fptype &dm3);

struct DecayInfo3 {
    fptype motherMass;
    fptype daug1Mass;
    fptype daug2Mass;
    fptype daug3Mass;
    fptype meson_radius;
    fptype mother_meson_radius;

    std::vector<ResonancePdf *> resonances;
};

/// DecayInfo3 plus the D0 lifetime and mixing parameters of the time-dependent
/// PDF (DalitzPlotHelpers.h:99-112 of the reference).
struct DecayInfo3t : public DecayInfo3 {
    Variable _tau;
    Variable _xmixing;
    Variable _ymixing;
    Variable _deltax;
    Variable _deltay;

    DecayInfo3t(Variable tau, Variable xmixing, Variable ymixing, Variable deltax, Variable deltay)
        : _tau(tau)
        , _xmixing(xmixing)
        , _ymixing(ymixing)
        , _deltax(deltax)
        , _deltay(deltay) {}
};

/// Cached (A, B) waves of one event for one resonance of the time-dependent PDF.
struct WaveHolder_s {
    fptype ai_real, ai_imag, bi_real, bi_imag;
};

}  // namespace GooFit
