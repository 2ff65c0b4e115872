// gf_device.h -- per-event FP64 arithmetic of the likelihood path, for gfx950.
//
// Each function states, in the same operation order, what the reference's
// device function computes (file:line given per function), so that results
// round the same way on both sides; the library is built with
// -ffp-contract=off so no multiply-add is fused behind our back.  The log of
// the likelihood metrics is this library's own (gf_math.h: within one unit in
// the last place of the correctly rounded value, half the device library's
// FP64 instructions); exp, erfc, pow, sqrt and the trigonometric functions are
// the device library's.
#pragma once

#include <hip/hip_runtime.h>

#include "gf_internal.h"
#include "gf_math.h"

namespace gfd {

// Wave-uniform read-only data (the program, parameter / normalisation /
// constant arrays, descriptor tables) is read through the constant address
// space: a uniform index then compiles to scalar loads through the scalar
// cache instead of per-lane vector loads followed by lane reads.  The data must
// not change while a kernel runs (it never does: it is uploaded between launches).
// The qualifier exists only in the device pass (the host pass has no address
// spaces and only parses these functions).
#if defined(__HIP_DEVICE_COMPILE__)
#define GF_CONST __attribute__((address_space(4)))
#else
#define GF_CONST
#endif
template <class T>
__device__ __forceinline__ const GF_CONST T *as_const(const T *p) {
    return (const GF_CONST T *)p;
}

// (re, im) += (cr + i ci) * (w.x + i w.y) as four fused multiply-adds (the code is
// built with contraction off, so the coherent sums ask for the fusion explicitly:
// half the instructions of separate products and sums, one rounding per step).
__device__ __forceinline__ void cmac(double &re, double &im, double cr, double ci, double2 w) {
    re = fma(cr, w.x, fma(-ci, w.y, re));
    im = fma(cr, w.y, fma(ci, w.x, im));
}

// The coherent sum sum_{i < nres} c_i W_i (device_DalitzPlot's totalAmp,
// src/PDFs/physics/Amp3Body.cu:88-108) with one fused multiply-add per real term, in
// groups of four resonances 4m .. 4m + 3: Re adds cr wr of the four, then -ci wi of
// the four; Im adds ci wr of the four, then cr wi of the four.  This is the order in
// which the FP64 matrix-core kernel adds its K terms (one 16x16x4 step takes one
// component of four resonances), so the vector and matrix-core sweeps give the same
// bits.  cpl(i) returns (cr, ci) of resonance i and is called for i < nres only; w
// holds NR >= nres waves, zero past nres.  A group that reaches past nres adds its
// terms past nres too (zero coupling, zero wave), as the matrix cores do: x + 0 is x,
// so only the sign of a zero sum could differ, which |A|^2 does not see.  Groups that
// lie wholly past nres are skipped, as the matrix-core kernel never forms them.
template <int NR, class Cpl>
__device__ __forceinline__ void coh_group(double &re, double &im, const double2 (&c)[4], const double2 *w) {
#pragma unroll
    for(int u = 0; u < 4; ++u) {
        re = fma(c[u].x, w[u].x, re);
        im = fma(c[u].y, w[u].x, im);
    }
#pragma unroll
    for(int u = 0; u < 4; ++u) {
        re = fma(-c[u].y, w[u].y, re);
        im = fma(c[u].x, w[u].y, im);
    }
}
template <int NR, class Cpl>
__device__ __forceinline__ void coh_sum(double &re, double &im, const Cpl &cpl, const double2 *w, int nres) {
    static_assert(NR % 4 == 0, "groups of four resonances");
#pragma unroll
    for(int m = 0; m < NR; m += 4) {
        double2 c[4];  // each coupling read once
        if(m + 4 <= nres) {
#pragma unroll
            for(int u = 0; u < 4; ++u)
                c[u] = cpl(m + u);
            coh_group<NR, Cpl>(re, im, c, w + m);
        } else if(m < nres) {
#pragma unroll
            for(int u = 0; u < 4; ++u)
                c[u] = m + u < nres ? cpl(m + u) : make_double2(0.0, 0.0);
            coh_group<NR, Cpl>(re, im, c, w + m);
        }
    }
}

// ------------------------------------------------------------------ grids

// Centre of local bin b of n equal bins on [lo, hi]: ((hi - lo) / n) * (b + 0.5) + lo,
// one rounding per operation in this order (MetricTaker.cu:81-84,
// SpecialResonanceIntegrator.cu:64-76).  Every grid kernel takes its cell centres
// from here, so the cells of an integral are the cells gf_grid_cells reports.
__host__ __device__ __forceinline__ double bin_centre(double lo, double hi, int n, int64_t b) {
    double x = hi - lo;
    x /= n;
    x *= ((double)b + 0.5);
    x += lo;
    return x;
}

// ---------------------------------------------------------------- leaf PDFs

// device_Gaussian, src/PDFs/basic/GaussianPdf.cu:7-17
__device__ __forceinline__ double gaussian(double x, double mean, double sigma) {
    return exp(-0.5 * (x - mean) * (x - mean) / (sigma * sigma));
}

// device_ExpGaus, src/PDFs/basic/ExpGausPdf.cu:6-24 (sqrt 2 truncated to 1.4142135623 as there)
__device__ __forceinline__ double expgaus(double x, double mean, double sigma, double alpha) {
    double ret    = 0.5 * alpha;
    double exparg = ret * (2 * mean + alpha * sigma * sigma - 2 * x);
    double erfarg = (mean + alpha * sigma * sigma - x) / (sigma * 1.4142135623);
    ret *= exp(exparg);
    ret *= erfc(erfarg);
    return ret;
}

// ------------------------------------------------------------ compensated sums

// Neumaier's variant of Kahan summation: s carries the rounded sum, c the
// accumulated rounding errors.
__device__ __forceinline__ void kb_add(double &s, double &c, double v) {
    double t = s + v;
    double d = (fabs(s) >= fabs(v)) ? ((s - t) + v) : ((v - t) + s);
    c += d;
    s = t;
}
__device__ __forceinline__ void kb_merge(double &s, double &c, double s2, double c2) {
    kb_add(s, c, s2);
    c += c2;
}
// The sum a compensated pair stands for.  Once s is infinite or NaN the
// correction is meaningless (inf - inf = NaN), and the plain sum's value -- the
// one the reference's uncompensated reduction gives -- is s itself.
__device__ __forceinline__ double kb_total(double s, double c) { return isfinite(s) ? s + c : s; }

// ------------------------------------------------------------------- metrics

// src/PDFs/MetricPointer.cu:8-58.  evt0/evt1 are the (bin entry, bin volume or
// error) of a binned event, unused for unbinned metrics.
__device__ __forceinline__ double metric(int m, double rawPdf, double evt0, double evt1, double norm,
                                         double totalEvents) {
    switch(m) {
    case 0: return rawPdf;
    case 1: rawPdf *= norm; return rawPdf > 0.0 ? -gfmath::log_d(rawPdf) : 0.0;
    case 2: return rawPdf * norm;
    case 3: {
        rawPdf *= norm;
        rawPdf *= evt1;
        if(rawPdf > 0) {
            double expEvents = totalEvents * rawPdf;
            return (expEvents - evt0 * gfmath::log_d(expEvents));
        }
        return 0;
    }
    case 4:
        rawPdf -= evt0;
        rawPdf /= evt1;
        rawPdf *= rawPdf;
        return rawPdf;
    case 5: {
        rawPdf *= norm;
        rawPdf *= evt1;
        double d = rawPdf * totalEvents - evt0;
        return d * d / (evt0 > 1 ? evt0 : 1);
    }
    default: return __builtin_nan("");
    }
}

// The NLL metric over a lane's events as one logarithm (gfmath::nll_add /
// nll_total): every kernel that reduces the NLL metric accumulates it this way, in
// its own event order per lane, in the registers of its (sum, residual) pair
// starting from (1, 0), so kernels that assign events to lanes alike give the same
// bits (the per-event values of evaluate_with_metric still use metric()).
__device__ __forceinline__ void nll_add(double &m, double &e, double rawPdf, double norm) {
    gfmath::nll_add(m, e, rawPdf * norm);
}
__device__ __forceinline__ double nll_total(double m, double e) { return gfmath::nll_total(m, e); }
constexpr int metric_nll = 1;

// --------------------------------------------------------- Dalitz kinematics

// inDalitz, src/PDFs/physics/DalitzPlotHelpers.cu:14-59
__device__ __forceinline__ bool in_dalitz(double m12, double m13, double bigM, double dm1, double dm2, double dm3) {
    double dm1pdm2  = dm1 + dm2;
    double bigMmdm3 = bigM - dm3;
    double bigMmdm1 = bigM - dm1;
    double m23      = bigM * bigM + dm1 * dm1 + dm2 * dm2 + dm3 * dm3 - m12 - m13;
    if(m23 < 0. || m23 > bigMmdm1 * bigMmdm1)
        return false;
    bool m12less     = (m12 < dm1pdm2 * dm1pdm2) ? false : true;
    bool m12grea     = (m12 > bigMmdm3 * bigMmdm3) ? false : true;
    double sqrtM12   = sqrt(m12);
    double dm11      = dm1 * dm1;
    double dm22      = dm2 * dm2;
    double dm33      = dm3 * dm3;
    double e1star    = 0.5 * (m12 - dm22 + dm11) / sqrtM12;
    double e3star    = 0.5 * (bigM * bigM - m12 - dm33) / sqrtM12;
    double rte1mdm11 = sqrt(e1star * e1star - dm11);
    double rte3mdm33 = sqrt(e3star * e3star - dm33);
    double minimum   = (e1star + e3star) * (e1star + e3star) - (rte1mdm11 + rte3mdm33) * (rte1mdm11 + rte3mdm33);
    bool m13less     = (m13 < minimum) ? false : true;
    double maximum   = (e1star + e3star) * (e1star + e3star) - (rte1mdm11 - rte3mdm33) * (rte1mdm11 - rte3mdm33);
    bool m13grea     = (m13 > maximum) ? false : true;
    return m12less && m12grea && m13less && m13grea;
}

// twoBodyCMmom, src/PDFs/physics/resonances/Resonance.cu:48-60
__device__ __forceinline__ double two_body_cm_mom(double rMassSq, double d1m, double d2m) {
    double kin1 = 1 - (d1m + d2m) * (d1m + d2m) / rMassSq;
    kin1        = kin1 >= 0 ? sqrt(kin1) : 1;
    double kin2 = 1 - (d1m - d2m) * (d1m - d2m) / rMassSq;
    kin2        = kin2 >= 0 ? sqrt(kin2) : 1;
    return 0.5 * sqrt(rMassSq) * kin1 * kin2;
}

// twoBodyCMMothermom, Resonance.cu:62-75
__device__ __forceinline__ double two_body_cm_mother_mom(double rMassSq, double dm, double d3m) {
    double kin1 = 1 - (dm + d3m) * (dm + d3m) / rMassSq;
    kin1        = kin1 >= 0 ? sqrt(kin1) : 1;
    double kin2 = 1 - (dm - d3m) * (dm - d3m) / rMassSq;
    kin2        = kin2 >= 0 ? sqrt(kin2) : 1;
    return 0.5 * rMassSq * kin1 * kin2 / dm;
}

// x^(2 spin + 1) for spin 0, 1, 2 by multiplication (the reference calls pow with
// the double exponent 2 spin + 1; products differ from pow by a few ulp, far
// inside the likelihood tolerance, and avoid FP64 pow's log/exp sequence).
__device__ __forceinline__ double odd_power(double x, int spin) {
    const double x2 = x * x;
    double r        = x;
    for(int k = 0; k < spin; ++k)
        r *= x2;
    return r;
}

// dampingFactorSquare, Resonance.cu:77-86 (spin-2 form as written there)
__device__ __forceinline__ double damping_factor_square(double cmmom, int spin, double mRadius) {
    double square  = mRadius * mRadius * cmmom * cmmom;
    double dfsq    = 2 * square;
    double dfsqres = 13 * square / ((square - 3) * (square - 3)) + 9 * square * square;
    return (spin == 2) ? dfsqres : dfsq;
}

// dampingFactorSquareNorm, Resonance.cu:88-97
__device__ __forceinline__ double damping_factor_square_norm(double cmmom, int spin, double mRadius) {
    double square  = mRadius * mRadius * cmmom * cmmom;
    double dfsq    = 1 + square;
    double dfsqres = dfsq + 8 + 2 * square + square * square;
    return (spin == 2) ? dfsqres : dfsq;
}

// spinFactor, Resonance.cu:99-155 (PAIR_12 = 0, PAIR_13 = 1, PAIR_23 = 2)
__device__ __forceinline__ double spin_factor(unsigned spin, const GfDecay &d, double m12, double m13, double m23,
                                              unsigned cyc) {
    if(0 == spin)
        return 1;
    double _mA  = (0 == cyc ? d.daug1Mass : (1 == cyc ? d.daug3Mass : d.daug2Mass));
    double _mB  = (0 == cyc ? d.daug2Mass : (1 == cyc ? d.daug1Mass : d.daug3Mass));
    double _mC  = (0 == cyc ? d.daug3Mass : (1 == cyc ? d.daug2Mass : d.daug1Mass));
    double _mAC = (0 == cyc ? m13 : (1 == cyc ? m23 : m12));
    double _mBC = (0 == cyc ? m23 : (1 == cyc ? m12 : m13));
    double _mAB = (0 == cyc ? m12 : (1 == cyc ? m13 : m23));
    double massFactor = 1.0 / (_mAB);
    double sFactor    = -1;
    sFactor *= ((_mBC - _mAC) + (massFactor * (d.motherMass * d.motherMass - _mC * _mC) * (_mA * _mA - _mB * _mB)));
    if(2 == spin) {
        sFactor *= sFactor;
        double t1        = d.motherMass * d.motherMass - _mC * _mC;
        double extraterm = ((_mAB - (2 * d.motherMass * d.motherMass) - (2 * _mC * _mC)) + massFactor * (t1 * t1));
        double t2        = _mA * _mA - _mB * _mB;
        extraterm *= ((_mAB - (2 * _mA * _mA) - (2 * _mB * _mB)) + massFactor * (t2 * t2));
        extraterm /= 3;
        sFactor -= extraterm;
    }
    return sFactor;
}

// plainBW<I>, src/PDFs/physics/resonances/RBW.cu:9-86
template <int I>
__device__ __forceinline__ double2 rbw(const GfDecay &d, double resmass, double reswidth, unsigned spin,
                                       unsigned cyclic_index, bool norm, double m12, double m13, double m23) {
    double2 result  = make_double2(0.0, 0.0);
    double resmass2 = resmass * resmass;
#pragma unroll
    for(int i = 0; i < I; i++) {
        double rMassSq      = (0 == cyclic_index ? m12 : (1 == cyclic_index ? m13 : m23));
        double mass_daug1   = 2 == cyclic_index ? d.daug2Mass : d.daug1Mass;
        double mass_daug2   = 0 == cyclic_index ? d.daug2Mass : d.daug3Mass;
        double bachelorMass = (0 == cyclic_index ? d.daug3Mass : (1 == cyclic_index ? d.daug2Mass : d.daug1Mass));
        double frFactor       = 1;
        double frFactorMother = 1;
        double measureDaughterMoms = two_body_cm_mom(rMassSq, mass_daug1, mass_daug2);
        double nominalDaughterMoms = two_body_cm_mom(resmass2, mass_daug1, mass_daug2);
        double measureDaughterMomsMother;
        double nominalDaughterMomsMother = 0;
        if(norm) {
            measureDaughterMomsMother = two_body_cm_mother_mom(rMassSq, d.motherMass, bachelorMass);
            nominalDaughterMomsMother = two_body_cm_mother_mom(resmass2, d.motherMass, bachelorMass);
        } else {
            measureDaughterMomsMother = two_body_cm_mom(d.motherMass * d.motherMass, sqrt(rMassSq), bachelorMass);
        }
        if(0 != spin) {
            if(norm) {
                frFactor = damping_factor_square_norm(nominalDaughterMoms, spin, d.meson_radius)
                           / damping_factor_square_norm(measureDaughterMoms, spin, d.meson_radius);
                frFactorMother = damping_factor_square_norm(nominalDaughterMomsMother, spin, d.mother_meson_radius)
                                 / damping_factor_square_norm(measureDaughterMomsMother, spin, d.mother_meson_radius);
            } else {
                frFactor       = damping_factor_square(measureDaughterMoms, spin, d.meson_radius);
                frFactorMother = damping_factor_square(measureDaughterMomsMother, spin, d.mother_meson_radius);
            }
        }
        double A  = (resmass2 - rMassSq);
        double B  = resmass2 * reswidth * odd_power(measureDaughterMoms / nominalDaughterMoms, spin) * frFactor
                   / sqrt(rMassSq);
        double C  = 1.0 / (A * A + B * B);
        double re = A * C, im = B * C;
        double s1 = sqrt(frFactor);
        re *= s1;
        im *= s1;
        double s2 = sqrt(frFactorMother);
        re *= s2;
        im *= s2;
        double sf = spin_factor(spin, d, m12, m13, m23, cyclic_index);
        re *= sf;
        im *= sf;
        result.x += re;
        result.y += im;
        if(I > 1)
            cyclic_index = (cyclic_index + 1) % 3;
    }
    return result;
}

// hFun, dh_dsFun, fsFun, dFun: Resonance.cu:5-46 (pi truncated to 3.14159265359 as there)
__device__ __forceinline__ double h_fun(double s, double daug2Mass, double daug3Mass) {
    const double _pi = 3.14159265359;
    double sm        = daug2Mass + daug3Mass;
    double sqrt_s    = sqrt(s);
    double k_s       = two_body_cm_mom(s, daug2Mass, daug3Mass);
    return ((2 / _pi) * (k_s / sqrt_s) * log((sqrt_s + 2 * k_s) / (sm)));
}
__device__ __forceinline__ double dh_ds_fun(double s, double daug2Mass, double daug3Mass) {
    const double _pi = 3.14159265359;
    double k_s       = two_body_cm_mom(s, daug2Mass, daug3Mass);
    return h_fun(s, daug2Mass, daug3Mass) * (1.0 / (8.0 * (k_s * k_s)) - 1.0 / (2.0 * s)) + 1.0 / (2.0 * _pi * s);
}
__device__ __forceinline__ double fs_fun(double s, double m2, double gam, double daug2Mass, double daug3Mass) {
    double k_s   = two_body_cm_mom(s, daug2Mass, daug3Mass);
    double k_Am2 = two_body_cm_mom(m2, daug2Mass, daug3Mass);
    double f     = gam * m2 / (k_Am2 * k_Am2 * k_Am2);
    f *= ((k_s * k_s) * (h_fun(s, daug2Mass, daug3Mass) - h_fun(m2, daug2Mass, daug3Mass))
          + (m2 - s) * (k_Am2 * k_Am2) * dh_ds_fun(m2, daug2Mass, daug3Mass));
    return f;
}
__device__ __forceinline__ double d_fun(double s, double daug2Mass, double daug3Mass) {
    const double _pi = 3.14159265359;
    double sm        = daug2Mass + daug3Mass;
    double sm24      = sm * sm / 4.0;
    double m         = sqrt(s);
    double k_m2      = two_body_cm_mom(s, daug2Mass, daug3Mass);
    return 3.0 / _pi * sm24 / (k_m2 * k_m2) * log((m + 2 * k_m2) / sm) + m / (2 * _pi * k_m2)
           - sm24 * m / (_pi * (k_m2 * k_m2 * k_m2));
}

// gouSak, src/PDFs/physics/resonances/GS.cu:9-75 (dFun/fsFun take daug2/daug3 masses in every channel, as there)
__device__ __forceinline__ double2 gounaris_sakurai(const GfDecay &d, double resmass, double reswidth, unsigned spin,
                                                   unsigned cyclic_index, bool norm, double m12, double m13,
                                                   double m23) {
    double rMassSq      = (0 == cyclic_index ? m12 : (1 == cyclic_index ? m13 : m23));
    double bachelorMass = (0 == cyclic_index ? d.daug3Mass : (1 == cyclic_index ? d.daug2Mass : d.daug1Mass));
    double frFactor     = 1;
    double frFactorD    = 1;
    resmass *= resmass;
    double md1 = (2 == cyclic_index ? d.daug2Mass : d.daug1Mass);
    double md2 = (0 == cyclic_index ? d.daug2Mass : d.daug3Mass);
    double measureDaughterMoms = two_body_cm_mom(rMassSq, md1, md2);
    double nominalDaughterMoms = two_body_cm_mom(resmass, md1, md2);
    double measureDaughterMomsMother;
    double nominalDaughterMomsMother = 0;
    if(norm) {
        measureDaughterMomsMother = two_body_cm_mother_mom(rMassSq, d.motherMass, bachelorMass);
        nominalDaughterMomsMother = two_body_cm_mother_mom(resmass, d.motherMass, bachelorMass);
    } else {
        measureDaughterMomsMother = two_body_cm_mom(d.motherMass * d.motherMass, sqrt(rMassSq), bachelorMass);
    }
    if(0 != spin) {
        if(norm) {
            frFactor = damping_factor_square_norm(nominalDaughterMoms, spin, d.meson_radius);
            frFactor /= damping_factor_square_norm(measureDaughterMoms, spin, d.meson_radius);
            frFactorD = damping_factor_square_norm(nominalDaughterMomsMother, spin, d.mother_meson_radius);
            frFactorD /= damping_factor_square_norm(measureDaughterMomsMother, spin, d.mother_meson_radius);
        } else {
            frFactor  = damping_factor_square(measureDaughterMoms, spin, d.meson_radius);
            frFactorD = damping_factor_square(measureDaughterMomsMother, spin, d.mother_meson_radius);
        }
    }
    double D = (1.0 + d_fun(resmass, d.daug2Mass, d.daug3Mass) * reswidth / sqrt(resmass));
    double E = resmass - rMassSq + fs_fun(rMassSq, resmass, reswidth, d.daug2Mass, d.daug3Mass);
    double F = sqrt(resmass) * reswidth * odd_power(measureDaughterMoms / nominalDaughterMoms, spin) * frFactor;
    D /= (E * E + F * F);
    double re = D * E, im = D * F;
    double s1 = sqrt(frFactor);
    re *= s1;
    im *= s1;
    double s2 = sqrt(frFactorD);
    re *= s2;
    im *= s2;
    double sf = spin_factor(spin, d, m12, m13, m23, cyclic_index);
    re *= sf;
    im *= sf;
    return make_double2(re, im);
}

// One resonance of a Dalitz node: getResonanceAmplitude (Amp3Body_TD.cu:36-39)
// dispatched on the wave-uniform kind instead of a function pointer.
__device__ __forceinline__ double2 lineshape(const GfResDesc &r, const double *__restrict__ P,
                                             const double *__restrict__ Cn, const GfDecay &d, double m12,
                                             double m13, double m23) {
    if(r.kind == GR_NONRES)
        return make_double2(1.0, 0.0);  // nonres, NonRes.cu:8-11
    double mass   = P[r.p];
    double width  = P[r.p + 1];
    unsigned spin = (unsigned)Cn[r.c];
    unsigned cyc  = (unsigned)Cn[r.c + 1];
    bool norm     = Cn[r.c + 2] != 0.0;
    switch(r.kind) {
    case GR_RBW: return rbw<1>(d, mass, width, spin, cyc, norm, m12, m13, m23);
    case GR_RBW_SYM: return rbw<2>(d, mass, width, spin, cyc, norm, m12, m13, m23);
    default: return gounaris_sakurai(d, mass, width, spin, cyc, norm, m12, m13, m23);
    }
}

// ------------------------------------------------------- time resolutions

// device_truth_resolution, src/PDFs/physics/TruthResolution.cu:6-27: the decay-time
// terms without smearing.
__device__ __forceinline__ double truth_resolution(double coshterm, double costerm, double sinhterm, double sinterm,
                                                   double tau, double dtime, double xmixing, double ymixing) {
    double ret = 0;
    dtime /= tau;
    ret += coshterm * cosh(ymixing * dtime);
    ret += costerm * cos(xmixing * dtime);
    ret -= 2 * sinhterm * sinh(ymixing * dtime);
    ret -= 2 * sinterm * sin(xmixing * dtime);
    ret *= exp(-dtime);
    return ret;
}

// One Gaussian of the three-Gaussian model (gaussian(), ThreeGaussResolution.cu:11-61):
// the decay-time terms convolved with a Gaussian of width adjSigma centred on adjTime;
// cos and sin are expanded to second and third order in x t / tau.
__device__ __forceinline__ void tg_gaussian(double &P1, double &P2, double &P3, double &P4, double tau, double adjTime,
                                            double xmixing, double ymixing, double adjSigma, double selbias) {
    const double sqrt_pi_over_2 = 1.0 / M_2_SQRTPI;
    const double inv_sqrt_2pi   = sqrt(0.5 * M_1_PI);
    const double sixth          = 1.0 / 6.0;
    const double a1  = adjSigma * M_SQRT2;  // 1 / sqrt(A)
    const double is  = 1 / adjSigma;
    const double a2  = 0.5 * a1;
    const double is2 = is * is;
    const double ng  = inv_sqrt_2pi * is;
    const double Cc  = 0.5 * adjTime * adjTime * is2;
    const double Bgn = -adjTime * is2;
    const double G   = 1 / tau;
    const double B   = G + Bgn;
    const double u0  = a2 * (B + selbias);
    const double u02 = u0 * u0;
    const double F   = a1 * exp(-Cc + u02);
    const double Ig0 = sqrt_pi_over_2 * erfc(u0);
    const double Ig1 = 0.5 * exp(-u02);
    const double Ig2 = Ig1 * u0 + 0.5 * Ig0;
    const double Ig3 = Ig1 * (u02 + 1);
    const double R   = xmixing * G * a1;
    const double R2  = R * R;
    const double It0 = F * Ig0;
    const double It1 = F * R * (Ig1 - u0 * Ig0);
    const double It2 = F * R2 * (Ig2 - u0 * Ig1 * 2 + u02 * Ig0);
    const double It3 = F * R2 * R * (Ig3 - u0 * Ig2 * 3 + u02 * Ig1 * 3 - u0 * u02 * Ig0);
    P2 = ng * (It0 - 0.5 * It2);
    P4 = ng * (It1 - sixth * It3);
    const double upy = a2 * (B - ymixing * G + selbias);
    const double umy = a2 * (B + ymixing * G + selbias);
    const double Ipy = a1 * exp(-Cc + upy * upy) * sqrt_pi_over_2 * erfc(upy);
    const double Imy = a1 * exp(-Cc + umy * umy) * sqrt_pi_over_2 * erfc(umy);
    P1 = ng * 0.5 * (Ipy + Imy);
    P3 = ng * 0.5 * (Ipy - Imy);
}

// The decay-time terms P1..P4 (cosh, cos, sinh, sin parts) of the three-Gaussian model,
// device_threegauss_resolution, ThreeGaussResolution.cu:63-115: core, tail and outlier
// Gaussians (fractions f_c, f_t, 1 - f_c - f_t) with per-event width sigma.  R = the 13
// parameters: fractions, biases and scales, selection bias, common offsets.
struct TimeTerms {
    double P1, P2, P3, P4;
};
__device__ __forceinline__ TimeTerms threegauss_terms(const GF_CONST double *__restrict__ R, double tau, double dtime,
                                                      double xmixing, double ymixing, double sigma) {
    const double fc = R[0], ft = R[1], fo = 1 - fc - ft;
    const double bc = R[2] + R[9], bt = R[4] + R[11], bo = R[6] + R[11];
    const double sc = R[3] * R[10], st = R[5] * R[12], so = R[7] * R[12];
    const double sb = R[8];
    double c1, c2, c3, c4, t1, t2, t3, t4, o1, o2, o3, o4;
    tg_gaussian(c1, c2, c3, c4, tau, dtime - bc * sigma, xmixing, ymixing, sc * sigma, sb);
    tg_gaussian(t1, t2, t3, t4, tau, dtime - bt * sigma, xmixing, ymixing, st * sigma, sb);
    tg_gaussian(o1, o2, o3, o4, tau, dtime - bo * sigma, xmixing, ymixing, so * sigma, sb);
    TimeTerms P;
    P.P1 = fc * c1 + ft * t1 + fo * o1;
    P.P2 = fc * c2 + ft * t2 + fo * o2;
    P.P3 = fc * c3 + ft * t3 + fo * o3;
    P.P4 = fc * c4 + ft * t4 + fo * o4;
    return P;
}
// ThreeGaussResolution.cu:117-122: the rate from the terms
__device__ __forceinline__ double threegauss_rate(const TimeTerms &P, double coshterm, double costerm,
                                                  double sinhterm, double sinterm) {
    double ret = 0;
    ret += coshterm * P.P1;
    ret += costerm * P.P2;
    ret -= 2 * sinhterm * P.P3;
    ret -= 2 * sinterm * P.P4;
    return ret;
}

// The resolution function of a time-dependent node, dispatched on its wave-uniform kind.
__device__ __forceinline__ double time_resolution(int kind, const GF_CONST double *__restrict__ R, double t1, double t2,
                                                  double abr, double abi, double tau, double dtime, double x,
                                                  double y, double sigma) {
    if(kind == GT_THREEGAUSS)
        return threegauss_rate(threegauss_terms(R, tau, dtime, x, y, sigma), t1, t2, abr, abi);
    return truth_resolution(t1, t2, abr, abi, tau, dtime, x, y);
}

// Counter-based random numbers: splitmix64 finaliser of (key, counter).  Any
// (key, counter) pair gives the same value on every run and device.
__host__ __device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
// uniform in [0, 1) with 53 random bits
__host__ __device__ __forceinline__ double unif01(uint64_t key, uint64_t ctr) {
    return (double)(mix64(key ^ mix64(ctr)) >> 11) * (1.0 / 9007199254740992.0);
}

}  // namespace gfd
