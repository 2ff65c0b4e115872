/*
 * goofit_oracle.c -- CPU restatement of GooFit 2.4.0's likelihood hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the HIP engine is
 * compared against; only the tests, the smoke check and the CPU-baseline leg
 * of the benchmark load it.  It is never part of the product path.
 *
 * It restates, function by function, the reference's device code as compiled
 * for its OpenMP Thrust backend (paths relative to the GooFit source tree):
 *   - the flattened-tree walk with ParameterContainer cursors
 *     (src/PDFs/ParameterContainer.cu:5-60, include/goofit/PDFs/ParameterContainer.h:9-47),
 *   - the basic and combining PDFs (src/PDFs/basic/, src/PDFs/combine/AddPdf.cu, ProdPdf.cu),
 *   - the metric functions (src/PDFs/MetricPointer.cu:8-58) and MetricTaker
 *     (src/PDFs/MetricTaker.cu:17-96),
 *   - the OpenMP reduction order of thrust::transform_reduce: omp_get_num_procs()
 *     contiguous chunks, each summed left to right from its first element, then
 *     init + chunk sums left to right (rocThrust omp/detail/reduce.inl:51-66,
 *     reduce_intervals.inl:72-88, default_decomposition.inl:50-52,
 *     system/detail/internal/decompose.h:67-97),
 *   - the three-body Dalitz amplitude: inDalitz (src/PDFs/physics/DalitzPlotHelpers.cu:14-59),
 *     lineshapes (src/PDFs/physics/resonances/{Resonance,RBW,GS,NonRes}.cu),
 *     the wave cache (detail/SpecialResonanceCalculator.cu:12-46), the grid
 *     integrals (detail/SpecialResonanceIntegrator.cu:7-121) and the coherent
 *     sum (src/PDFs/physics/Amp3Body.cu:60-122),
 *   - the time-dependent Dalitz amplitude: the rate with mistag and charm tag
 *     (src/PDFs/physics/Amp3Body_TD.cu:41-257), the truth and three-Gaussian
 *     resolutions and their normalisations (TruthResolution.cu, ThreeGaussResolution.cu),
 *     the A/B waves (detail/SpecialWaveCalculator.cu:13-63), the ThreeComplex grid
 *     integrals (detail/SpecialDalitzIntegrator.cu:9-126) and the host assembly of
 *     the normalisation (Amp3Body_TD.cu:686-751).
 *
 * Built with -ffp-contract=off so that every product and sum rounds exactly
 * as written (no fused multiply-add), as on the reference's x86 OMP build.
 *
 * Pinning (what fixes each function without trusting this restatement):
 *   Gaussian, Exp, Add, Prod, the 1-D grid and the metrics: the reference's
 *   known-answer tests (tests/simple/NormalizeTest.cpp, AddNormTest.cpp,
 *   SimpleTest.cpp) and closed forms; ExpGaus: the exponentially modified
 *   Gaussian closed form (scipy.stats.exponnorm); inDalitz: brute-force
 *   three-body kinematics; twoBodyCMmom: PDG (eq. 49.17 in recent editions);
 *   RBW/GS: independent Python formulas of the published lineshapes plus
 *   the quirks A5/A6 reproduced on purpose; integrals: Hermitian and positive
 *   semi-definite, and the closed-form area for a flat amplitude; the chunked
 *   summation: compared with exact (fsum) sums.  Time-dependent Dalitz: decay-time
 *   integrals in closed form, the resolutions' normalisations against numerical
 *   integrals of their event functions, the narrow-resolution limit, the x = y = 0
 *   reduction to the Amp3Body density, the charm-tag and mistag identities, and the
 *   normalisation as the integral of the event density (tests/test_tddp_oracle.py).
 *   RBW spin 1 and 2 away from the pole: the normalised Blatt-Weisskopf ratios
 *   against the spherical Hankel functions (scipy), with the mother-side momentum
 *   s / (2 M) that twoBodyCMMothermom's clamped square roots give inside the plot;
 *   GS away from its pole: the published Gounaris-Sakurai f(s) with h' by
 *   complex-step differentiation and the width term as GS.cu:64 writes it.
 *   Polynomial, OffsetPolynomial, ExpOffset, ExpPoly, ExpPolyOffset: numpy
 *   polynomial closed forms of the parameter layouts the reference's constructors
 *   register; MultiPolynomial: unit-coefficient monomials in the coefficient order
 *   the reference documents (PolynomialPdf.h:22-33) and a brute-force monomial sum in
 *   three observables; the extended AddPdf: the extended NLL from scipy densities;
 *   RBW_Sym: the sum of its two channels recomputed from four-vectors; RBW and GS
 *   with unnormalised barriers (norm = false, including A5's spin-2 form): the
 *   lineshapes recomputed from four-vector momenta and Zemach factors; GS in the
 *   pi0 pi+- channels (A6): the published functions of two charged pions with the
 *   pair's own momenta in the width term (tests/test_oracle_pins_more.py).
 *   Parity unpinned: the summation order inside MultiPolynomial (the terms are the
 *   documented ones; their order of addition is the restatement's own).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/goofit_hip.h"

#define GFO_MAX_CACHE (16 * 20) /* cResonances[16 * 20], src/PDFs/physics/Amp3Body.cu:54 */

typedef struct {
    double re, im;
} cplx;

static inline cplx cmake(double re, double im) {
    cplx c = {re, im};
    return c;
}
/* thrust::complex operator* (lhs.re*rhs.re - lhs.im*rhs.im, lhs.re*rhs.im + lhs.im*rhs.re) */
static inline cplx cmul(cplx a, cplx b) {
    return cmake(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
static inline cplx cconj(cplx a) { return cmake(a.re, -a.im); }
static inline cplx cadd(cplx a, cplx b) { return cmake(a.re + b.re, a.im + b.im); }
static inline cplx cscale(cplx a, double s) { return cmake(a.re * s, a.im * s); }

typedef struct gfo_prog {
    int *funcs;
    int nfunc;
    double *params;
    int np;
    double *consts;
    int nc;
    double *obs;
    int no;
    double *norms;
    int nn;
    /* c_motherMass .. c_mother_meson_radius (src/PDFs/detail/Globals.cu:27-33) */
    double motherMass, daug1Mass, daug2Mass, daug3Mass, meson_radius, mother_meson_radius;
    /* cResonances[16 * cacheToUse + i][evtNum] as interleaved (re, im) */
    const double *cache[GFO_MAX_CACHE];
    int64_t cache_len[GFO_MAX_CACHE];
    /* Amp3Body_TD: cWaves[16 * cacheToUse + i][evtNum] holds (ai, bi); ai above, bi here */
    const double *cache_b[GFO_MAX_CACHE];
    double total_events; /* c_totalEvents (src/PDFs/PdfBase.cu:438) */
    int error;           /* set when an evaluation walks outside the arrays */
} gfo_prog;

/* ParameterContainer (include/goofit/PDFs/ParameterContainer.h:9-47) */
typedef struct {
    int parameterIdx, constantIdx, observableIdx, normalIdx, funcIdx;
} pcur;

static inline double P(const gfo_prog *g, const pcur *pc, int i) { return g->params[pc->parameterIdx + i + 1]; }
static inline double C(const gfo_prog *g, const pcur *pc, int i) { return g->consts[pc->constantIdx + i + 1]; }
static inline double O(const gfo_prog *g, const pcur *pc, int i) { return g->obs[pc->observableIdx + i + 1]; }
static inline double N(const gfo_prog *g, const pcur *pc, int i) { return g->norms[pc->normalIdx + i + 1]; }
static inline int NP(const gfo_prog *g, const pcur *pc) { return (int)g->params[pc->parameterIdx]; }
static inline int NC(const gfo_prog *g, const pcur *pc) { return (int)g->consts[pc->constantIdx]; }
static inline int NO(const gfo_prog *g, const pcur *pc) { return (int)g->obs[pc->observableIdx]; }

/* incrementIndex(funcs, params, cons, obs, norms), ParameterContainer.cu:20-46 */
static inline void inc5(pcur *pc, int f, int p, int c, int o, int n) {
    pc->funcIdx += f;
    pc->parameterIdx += p + 1;
    pc->constantIdx += c + 1;
    pc->observableIdx += o + 1;
    pc->normalIdx += n + 1;
}
/* slow incrementIndex(), ParameterContainer.cu:48-60 */
static inline void inc(const gfo_prog *g, pcur *pc) {
    int np = (int)g->params[pc->parameterIdx];
    int nc = (int)g->consts[pc->constantIdx];
    int no = (int)g->obs[pc->observableIdx];
    int nn = (int)g->norms[pc->normalIdx];
    pc->funcIdx++;
    pc->parameterIdx += np + 1;
    pc->constantIdx += nc + 1;
    pc->observableIdx += no + 1;
    pc->normalIdx += nn + 1;
}

static double call_function(gfo_prog *g, const double *evt, pcur *pc);

/* ---------------------------------------------------------------- basic PDFs */

/* device_Gaussian, src/PDFs/basic/GaussianPdf.cu:7-17 */
static double f_gaussian(gfo_prog *g, const double *evt, pcur *pc) {
    int id       = (int)O(g, pc, 0);
    double x     = evt[id];
    double mean  = P(g, pc, 0);
    double sigma = P(g, pc, 1);
    inc5(pc, 1, 2, 0, 1, 1);
    return exp(-0.5 * (x - mean) * (x - mean) / (sigma * sigma));
}

/* device_Exp, src/PDFs/basic/ExpPdf.cu:7-17 */
static double f_exp(gfo_prog *g, const double *evt, pcur *pc) {
    int id       = (int)O(g, pc, 0);
    double alpha = P(g, pc, 0);
    double x     = evt[id];
    double ret   = exp(alpha * x);
    inc5(pc, 1, 1, 0, 1, 1);
    return ret;
}

/* device_ExpOffset, src/PDFs/basic/ExpPdf.cu:19-30 */
static double f_exp_offset(gfo_prog *g, const double *evt, pcur *pc) {
    int id   = (int)O(g, pc, 0);
    double x = evt[id];
    x -= P(g, pc, 0);
    double alpha = P(g, pc, 1);
    double ret   = exp(alpha * x);
    inc5(pc, 1, 2, 0, 1, 1);
    return ret;
}

/* device_ExpPoly, src/PDFs/basic/ExpPdf.cu:32-48 */
static double f_exp_poly(gfo_prog *g, const double *evt, pcur *pc) {
    int id        = (int)O(g, pc, 0);
    double x      = evt[id];
    double exparg = 0;
    int np        = NP(g, pc);
    for(int i = 0; i < np; ++i)
        exparg += pow(x, i) * P(g, pc, i);
    double ret = exp(exparg);
    inc5(pc, 1, np, 0, 1, 1);
    return ret;
}

/* device_ExpPolyOffset, src/PDFs/basic/ExpPdf.cu:50-67 */
static double f_exp_poly_offset(gfo_prog *g, const double *evt, pcur *pc) {
    int id   = (int)O(g, pc, 0);
    double x = evt[id];
    x -= P(g, pc, 0);
    double exparg = 0;
    int np        = NP(g, pc);
    for(int i = 1; i < np; ++i)
        exparg += pow(x, i) * P(g, pc, i);
    double ret = exp(exparg);
    inc5(pc, 1, np, 0, 1, 1);
    return ret;
}

/* device_ExpGaus, src/PDFs/basic/ExpGausPdf.cu:6-24 (sqrt(2) truncated to 1.4142135623 as there) */
static double f_expgaus(gfo_prog *g, const double *evt, pcur *pc) {
    int id        = (int)O(g, pc, 0);
    double x      = evt[id];
    double mean   = P(g, pc, 0);
    double sigma  = P(g, pc, 1);
    double alpha  = P(g, pc, 2);
    double ret    = 0.5 * alpha;
    double exparg = ret * (2 * mean + alpha * sigma * sigma - 2 * x);
    double erfarg = (mean + alpha * sigma * sigma - x) / (sigma * 1.4142135623);
    ret *= exp(exparg);
    ret *= erfc(erfarg);
    inc5(pc, 1, 3, 0, 1, 1);
    return ret;
}

/* device_Polynomial, src/PDFs/basic/PolynomialPdf.cu:8-25 */
static double f_polynomial(gfo_prog *g, const double *evt, pcur *pc) {
    int id           = (int)O(g, pc, 0);
    int numParams    = NP(g, pc);
    int lowestDegree = (int)C(g, pc, 0);
    double x         = evt[id];
    double ret       = 0;
    for(int i = 0; i < numParams; ++i)
        ret += P(g, pc, i) * pow(x, lowestDegree + i);
    inc5(pc, 1, numParams, 1, 1, 1);
    return ret;
}

/* device_OffsetPolynomial, src/PDFs/basic/PolynomialPdf.cu:28-45 */
static double f_offset_polynomial(gfo_prog *g, const double *evt, pcur *pc) {
    int id           = (int)O(g, pc, 0);
    int numParams    = NP(g, pc);
    int lowestDegree = (int)C(g, pc, 0);
    double x         = evt[id];
    (void)x;
    double ret = 0;
    for(int i = 2; i < numParams; ++i)
        ret += P(g, pc, i) * pow(x, lowestDegree + i - 2);
    inc5(pc, 1, numParams, 1, 1, 1);
    return ret;
}

/* device_MultiPolynomial, src/PDFs/basic/PolynomialPdf.cu:47-118 */
static double f_multi_polynomial(gfo_prog *g, const double *evt, pcur *pc) {
    int num_constants   = NC(g, pc);
    int num_parameters  = NP(g, pc);
    int num_observables = NO(g, pc);
    int maxDegree       = (int)C(g, pc, 0) + 1;
    int numBoxes        = 1;
    for(int i = 0; i < num_observables; ++i)
        numBoxes *= maxDegree;
    int coeffNumber = num_observables;
    double ret      = P(g, pc, coeffNumber);
    coeffNumber++;
    for(int i = 1; i < numBoxes; ++i) {
        double currTerm  = 1;
        int currIndex    = i;
        int sumOfIndices = 0;
        for(int j = 0; j < num_observables; ++j) {
            int id        = (int)O(g, pc, j);
            double x      = evt[id];
            double offset = P(g, pc, j);
            x -= offset;
            int currPower = currIndex % maxDegree;
            currIndex /= maxDegree;
            currTerm *= pow(x, currPower);
            sumOfIndices += currPower;
        }
        if(sumOfIndices >= maxDegree)
            continue;
        double coefficient = P(g, pc, coeffNumber);
        coeffNumber++;
        currTerm *= coefficient;
        ret += currTerm;
    }
    inc5(pc, 1, num_parameters, num_constants, num_observables, 1);
    return ret;
}

/* ------------------------------------------------------------ combining PDFs */

/* device_AddPdfs, src/PDFs/combine/AddPdf.cu:16-51 */
static double f_add(gfo_prog *g, const double *evt, pcur *pc) {
    int numParameters  = NP(g, pc);
    double ret         = 0;
    double totalWeight = 0;
    pcur local         = *pc;
    inc(g, pc);
    for(int i = 0; i < numParameters; i++) {
        double weight = P(g, &local, i);
        totalWeight += weight;
        double norm = N(g, pc, 0);
        double curr = call_function(g, evt, pc);
        ret += weight * curr * norm;
    }
    double normFactor = N(g, pc, 0);
    double last       = call_function(g, evt, pc);
    ret += (1 - totalWeight) * last * normFactor;
    return ret;
}

/* device_AddPdfsExt, src/PDFs/combine/AddPdf.cu:53-79 */
static double f_add_ext(gfo_prog *g, const double *evt, pcur *pc) {
    int numParameters  = NP(g, pc);
    double ret         = 0;
    double totalWeight = 0;
    pcur local         = *pc;
    inc(g, pc);
    for(int i = 0; i < numParameters; i++) {
        double weight     = P(g, &local, i);
        double normFactor = N(g, pc, 0);
        double curr       = call_function(g, evt, pc);
        ret += weight * curr * normFactor;
        totalWeight += weight;
    }
    ret /= totalWeight;
    return ret;
}

/* device_ProdPdfs, src/PDFs/combine/ProdPdf.cu:9-26 */
static double f_prod(gfo_prog *g, const double *evt, pcur *pc) {
    int numCons  = NC(g, pc);
    int numComps = (int)C(g, pc, 0);
    int numObs   = NO(g, pc);
    double ret   = 1;
    inc5(pc, 1, 0, numCons, numObs, 1);
    for(int i = 0; i < numComps; i++) {
        double norm = N(g, pc, 0);
        double curr = call_function(g, evt, pc);
        curr *= norm;
        ret *= curr;
    }
    return ret;
}

/* ---------------------------------------------------------- Dalitz kinematics */

/* inDalitz, src/PDFs/physics/DalitzPlotHelpers.cu:14-59 */
int gfo_in_dalitz(double m12, double m13, double bigM, double dm1, double dm2, double dm3) {
    double dm1pdm2  = dm1 + dm2;
    double bigMmdm3 = bigM - dm3;
    double bigMmdm1 = bigM - dm1;
    double m23      = bigM * bigM + dm1 * dm1 + dm2 * dm2 + dm3 * dm3 - m12 - m13;
    if(m23 < 0. || m23 > bigMmdm1 * bigMmdm1)
        return 0;
    int m12less    = (m12 < dm1pdm2 * dm1pdm2) ? 0 : 1;
    int m12grea    = (m12 > bigMmdm3 * bigMmdm3) ? 0 : 1;
    double sqrtM12 = sqrt(m12);
    double dm11    = dm1 * dm1;
    double dm22    = dm2 * dm2;
    double dm33    = dm3 * dm3;
    double e1star  = 0.5 * (m12 - dm22 + dm11) / sqrtM12;
    double e3star  = 0.5 * (bigM * bigM - m12 - dm33) / sqrtM12;
    double rte1mdm11 = sqrt(e1star * e1star - dm11);
    double rte3mdm33 = sqrt(e3star * e3star - dm33);
    double minimum   = (e1star + e3star) * (e1star + e3star) - (rte1mdm11 + rte3mdm33) * (rte1mdm11 + rte3mdm33);
    int m13less      = (m13 < minimum) ? 0 : 1;
    double maximum   = (e1star + e3star) * (e1star + e3star) - (rte1mdm11 - rte3mdm33) * (rte1mdm11 - rte3mdm33);
    int m13grea      = (m13 > maximum) ? 0 : 1;
    return m12less && m12grea && m13less && m13grea;
}

/* twoBodyCMmom, src/PDFs/physics/resonances/Resonance.cu:48-60 */
double gfo_two_body_cm_mom(double rMassSq, double d1m, double d2m) {
    double kin1 = 1 - (d1m + d2m) * (d1m + d2m) / rMassSq;
    kin1        = kin1 >= 0 ? sqrt(kin1) : 1;
    double kin2 = 1 - (d1m - d2m) * (d1m - d2m) / rMassSq;
    kin2        = kin2 >= 0 ? sqrt(kin2) : 1;
    return 0.5 * sqrt(rMassSq) * kin1 * kin2;
}

/* twoBodyCMMothermom, Resonance.cu:62-75 */
static double two_body_cm_mother_mom(double rMassSq, double dm, double d3m) {
    double kin1 = 1 - (dm + d3m) * (dm + d3m) / rMassSq;
    if(kin1 >= 0)
        kin1 = sqrt(kin1);
    else
        kin1 = 1;
    double kin2 = 1 - (dm - d3m) * (dm - d3m) / rMassSq;
    if(kin2 >= 0)
        kin2 = sqrt(kin2);
    else
        kin2 = 1;
    return 0.5 * rMassSq * kin1 * kin2 / dm;
}

/* dampingFactorSquare, Resonance.cu:77-86 (spin-2 form reproduced as written) */
static double damping_factor_square(double cmmom, int spin, double mRadius) {
    double square  = mRadius * mRadius * cmmom * cmmom;
    double dfsq    = 2 * square;
    double dfsqres = 13 * square / pow(square - 3, 2) + 9 * square * square;
    return (spin == 2) ? dfsqres : dfsq;
}

/* dampingFactorSquareNorm, Resonance.cu:88-97 */
static double damping_factor_square_norm(double cmmom, int spin, double mRadius) {
    double square  = mRadius * mRadius * cmmom * cmmom;
    double dfsq    = 1 + square;
    double dfsqres = dfsq + 8 + 2 * square + square * square;
    return (spin == 2) ? dfsqres : dfsq;
}

/* spinFactor, Resonance.cu:99-155 (PAIR_12 = 0, PAIR_13 = 1, PAIR_23 = 2) */
static double spin_factor(unsigned spin, double motherMass, double daug1Mass, double daug2Mass, double daug3Mass,
                          double m12, double m13, double m23, unsigned cyclic_index) {
    if(0 == spin)
        return 1;
    double _mA  = (0 == cyclic_index ? daug1Mass : (1 == cyclic_index ? daug3Mass : daug2Mass));
    double _mB  = (0 == cyclic_index ? daug2Mass : (1 == cyclic_index ? daug1Mass : daug3Mass));
    double _mC  = (0 == cyclic_index ? daug3Mass : (1 == cyclic_index ? daug2Mass : daug1Mass));
    double _mAC = (0 == cyclic_index ? m13 : (1 == cyclic_index ? m23 : m12));
    double _mBC = (0 == cyclic_index ? m23 : (1 == cyclic_index ? m12 : m13));
    double _mAB = (0 == cyclic_index ? m12 : (1 == cyclic_index ? m13 : m23));
    double massFactor = 1.0 / (_mAB);
    double sFactor    = -1;
    sFactor *= ((_mBC - _mAC) + (massFactor * (motherMass * motherMass - _mC * _mC) * (_mA * _mA - _mB * _mB)));
    if(2 == spin) {
        sFactor *= sFactor;
        double t1 = motherMass * motherMass - _mC * _mC;
        double extraterm = ((_mAB - (2 * motherMass * motherMass) - (2 * _mC * _mC)) + massFactor * (t1 * t1));
        double t2 = _mA * _mA - _mB * _mB;
        extraterm *= ((_mAB - (2 * _mA * _mA) - (2 * _mB * _mB)) + massFactor * (t2 * t2));
        extraterm /= 3;
        sFactor -= extraterm;
    }
    return sFactor;
}

/* plainBW<I>, src/PDFs/physics/resonances/RBW.cu:9-86 */
static cplx f_rbw(gfo_prog *g, int I, double m12, double m13, double m23, pcur *pc) {
    unsigned spin         = (unsigned)C(g, pc, 0);
    unsigned cyclic_index = (unsigned)C(g, pc, 1);
    int norm              = C(g, pc, 2) != 0.0;
    double resmass        = P(g, pc, 0);
    double reswidth       = P(g, pc, 1);
    cplx result           = cmake(0.0, 0.0);
    double resmass2       = resmass * resmass;
    for(int i = 0; i < I; i++) {
        double rMassSq    = (0 == cyclic_index ? m12 : (1 == cyclic_index ? m13 : m23));
        double mass_daug1 = 2 == cyclic_index ? g->daug2Mass : g->daug1Mass;
        double mass_daug2 = 0 == cyclic_index ? g->daug2Mass : g->daug3Mass;
        double bachelorMass
            = (0 == cyclic_index ? g->daug3Mass : (1 == cyclic_index ? g->daug2Mass : g->daug1Mass));
        double frFactor            = 1;
        double frFactorMother      = 1;
        double measureDaughterMoms = gfo_two_body_cm_mom(rMassSq, mass_daug1, mass_daug2);
        double nominalDaughterMoms = gfo_two_body_cm_mom(resmass2, mass_daug1, mass_daug2);
        double measureDaughterMomsMother = 0;
        double nominalDaughterMomsMother = 0;
        if(norm) {
            measureDaughterMomsMother = two_body_cm_mother_mom(rMassSq, g->motherMass, bachelorMass);
            nominalDaughterMomsMother = two_body_cm_mother_mom(resmass2, g->motherMass, bachelorMass);
        } else {
            measureDaughterMomsMother = gfo_two_body_cm_mom(g->motherMass * g->motherMass, sqrt(rMassSq), bachelorMass);
        }
        if(0 != spin) {
            if(norm) {
                frFactor = damping_factor_square_norm(nominalDaughterMoms, spin, g->meson_radius)
                           / damping_factor_square_norm(measureDaughterMoms, spin, g->meson_radius);
                frFactorMother
                    = damping_factor_square_norm(nominalDaughterMomsMother, spin, g->mother_meson_radius)
                      / damping_factor_square_norm(measureDaughterMomsMother, spin, g->mother_meson_radius);
            } else {
                frFactor       = damping_factor_square(measureDaughterMoms, spin, g->meson_radius);
                frFactorMother = damping_factor_square(measureDaughterMomsMother, spin, g->mother_meson_radius);
            }
        }
        double A = (resmass2 - rMassSq);
        double B = resmass2 * reswidth * pow(measureDaughterMoms / nominalDaughterMoms, 2.0 * spin + 1) * frFactor
                   / sqrt(rMassSq);
        double Cc = 1.0 / (A * A + B * B);
        cplx ret  = cmake(A * Cc, B * Cc);
        ret       = cscale(ret, sqrt(frFactor));
        ret       = cscale(ret, sqrt(frFactorMother));
        ret       = cscale(ret, spin_factor(spin, g->motherMass, g->daug1Mass, g->daug2Mass, g->daug3Mass, m12, m13,
                                            m23, cyclic_index));
        result    = cadd(result, ret);
        if(I > 1)
            cyclic_index = (cyclic_index + 1) % 3;
    }
    inc5(pc, 1, 2, 3, 0, 1);
    return result;
}

/* hFun, dh_dsFun, fsFun, dFun: Resonance.cu:5-46 (pi truncated to 3.14159265359 as there) */
static double h_fun(double s, double daug2Mass, double daug3Mass) {
    const double _pi = 3.14159265359;
    double sm        = daug2Mass + daug3Mass;
    double sqrt_s    = sqrt(s);
    double k_s       = gfo_two_body_cm_mom(s, daug2Mass, daug3Mass);
    return ((2 / _pi) * (k_s / sqrt_s) * log((sqrt_s + 2 * k_s) / (sm)));
}
static double dh_ds_fun(double s, double daug2Mass, double daug3Mass) {
    const double _pi = 3.14159265359;
    double k_s       = gfo_two_body_cm_mom(s, daug2Mass, daug3Mass);
    return h_fun(s, daug2Mass, daug3Mass) * (1.0 / (8.0 * (k_s * k_s)) - 1.0 / (2.0 * s)) + 1.0 / (2.0 * _pi * s);
}
static double fs_fun(double s, double m2, double gam, double daug2Mass, double daug3Mass) {
    double k_s   = gfo_two_body_cm_mom(s, daug2Mass, daug3Mass);
    double k_Am2 = gfo_two_body_cm_mom(m2, daug2Mass, daug3Mass);
    double f     = gam * m2 / (k_Am2 * k_Am2 * k_Am2);
    f *= ((k_s * k_s) * (h_fun(s, daug2Mass, daug3Mass) - h_fun(m2, daug2Mass, daug3Mass))
          + (m2 - s) * (k_Am2 * k_Am2) * dh_ds_fun(m2, daug2Mass, daug3Mass));
    return f;
}
static double d_fun(double s, double daug2Mass, double daug3Mass) {
    const double _pi = 3.14159265359;
    double sm        = daug2Mass + daug3Mass;
    double sm24      = sm * sm / 4.0;
    double m         = sqrt(s);
    double k_m2      = gfo_two_body_cm_mom(s, daug2Mass, daug3Mass);
    return 3.0 / _pi * sm24 / (k_m2 * k_m2) * log((m + 2 * k_m2) / sm) + m / (2 * _pi * k_m2)
           - sm24 * m / (_pi * (k_m2 * k_m2 * k_m2));
}

/* gouSak, src/PDFs/physics/resonances/GS.cu:9-75 (dFun/fsFun use daug2/daug3 masses in every channel, as there) */
static cplx f_gs(gfo_prog *g, double m12, double m13, double m23, pcur *pc) {
    unsigned spin         = (unsigned)C(g, pc, 0);
    unsigned cyclic_index = (unsigned)C(g, pc, 1);
    int norm              = C(g, pc, 2) != 0.0;
    double resmass        = P(g, pc, 0);
    double reswidth       = P(g, pc, 1);
    double rMassSq        = (0 == cyclic_index ? m12 : (1 == cyclic_index ? m13 : m23));
    double bachelorMass   = (0 == cyclic_index ? g->daug3Mass : (1 == cyclic_index ? g->daug2Mass : g->daug1Mass));
    double frFactor       = 1;
    double frFactorD      = 1;
    resmass *= resmass;
    double measureDaughterMoms = gfo_two_body_cm_mom(rMassSq, (2 == cyclic_index ? g->daug2Mass : g->daug1Mass),
                                                     (0 == cyclic_index ? g->daug2Mass : g->daug3Mass));
    double nominalDaughterMoms = gfo_two_body_cm_mom(resmass, (2 == cyclic_index ? g->daug2Mass : g->daug1Mass),
                                                     (0 == cyclic_index ? g->daug2Mass : g->daug3Mass));
    double measureDaughterMomsMother = 0;
    double nominalDaughterMomsMother = 0;
    if(norm) {
        measureDaughterMomsMother = two_body_cm_mother_mom(rMassSq, g->motherMass, bachelorMass);
        nominalDaughterMomsMother = two_body_cm_mother_mom(resmass, g->motherMass, bachelorMass);
    } else {
        measureDaughterMomsMother = gfo_two_body_cm_mom(g->motherMass * g->motherMass, sqrt(rMassSq), bachelorMass);
    }
    if(0 != spin) {
        if(norm) {
            frFactor = damping_factor_square_norm(nominalDaughterMoms, spin, g->meson_radius);
            frFactor /= damping_factor_square_norm(measureDaughterMoms, spin, g->meson_radius);
            frFactorD = damping_factor_square_norm(nominalDaughterMomsMother, spin, g->mother_meson_radius);
            frFactorD /= damping_factor_square_norm(measureDaughterMomsMother, spin, g->mother_meson_radius);
        } else {
            frFactor  = damping_factor_square(measureDaughterMoms, spin, g->meson_radius);
            frFactorD = damping_factor_square(measureDaughterMomsMother, spin, g->mother_meson_radius);
        }
    }
    double D = (1.0 + d_fun(resmass, g->daug2Mass, g->daug3Mass) * reswidth / sqrt(resmass));
    double E = resmass - rMassSq + fs_fun(rMassSq, resmass, reswidth, g->daug2Mass, g->daug3Mass);
    double F = sqrt(resmass) * reswidth * pow(measureDaughterMoms / nominalDaughterMoms, 2.0 * spin + 1) * frFactor;
    D /= (E * E + F * F);
    cplx retur = cmake(D * E, D * F);
    retur      = cscale(retur, sqrt(frFactor));
    retur      = cscale(retur, sqrt(frFactorD));
    retur      = cscale(retur, spin_factor(spin, g->motherMass, g->daug1Mass, g->daug2Mass, g->daug3Mass, m12, m13,
                                           m23, cyclic_index));
    inc5(pc, 1, 2, 3, 0, 1);
    return retur;
}

/* nonres, src/PDFs/physics/resonances/NonRes.cu:8-11 */
static cplx f_nonres(pcur *pc) {
    inc5(pc, 1, 0, 0, 0, 1);
    return cmake(1.0, 0.0);
}

/* getResonanceAmplitude, src/PDFs/physics/Amp3Body_TD.cu:36-39 */
static cplx resonance_amplitude(gfo_prog *g, double m12, double m13, double m23, pcur *pc) {
    switch(g->funcs[pc->funcIdx]) {
    case GF_FN_RBW: return f_rbw(g, 1, m12, m13, m23, pc);
    case GF_FN_RBW_SYM: return f_rbw(g, 2, m12, m13, m23, pc);
    case GF_FN_GS: return f_gs(g, m12, m13, m23, pc);
    case GF_FN_NONRES: return f_nonres(pc);
    default: g->error = 1; return cmake(0, 0);
    }
}

/* device_DalitzPlot, src/PDFs/physics/Amp3Body.cu:60-122 */
static double f_dalitz(gfo_prog *g, const double *evt, pcur *pc) {
    int num_obs           = NO(g, pc);
    int id_m12            = (int)O(g, pc, 0);
    int id_m13            = (int)O(g, pc, 1);
    int id_num            = (int)O(g, pc, 2);
    double m12            = evt[id_m12];
    double m13            = evt[id_m13];
    unsigned numResonances = (unsigned)C(g, pc, 0);
    unsigned cacheToUse   = (unsigned)C(g, pc, 1);
    if(!gfo_in_dalitz(m12, m13, g->motherMass, g->daug1Mass, g->daug2Mass, g->daug3Mass)) {
        inc5(pc, 1, (int)numResonances * 2, 2, num_obs, 1);
        for(unsigned i = 0; i < numResonances; i++)
            inc(g, pc);
        inc(g, pc);
        return 0;
    }
    double evtIndex = evt[id_num];
    int evtNum      = (int)floor(0.5 + evtIndex);
    cplx totalAmp   = cmake(0, 0);
    for(unsigned i = 0; i < numResonances; ++i) {
        cplx amp        = cmake(P(g, pc, (int)i * 2), P(g, pc, (int)i * 2 + 1));
        unsigned slot   = i + 16 * cacheToUse;
        const double *w = slot < GFO_MAX_CACHE ? g->cache[slot] : NULL;
        if(!w || evtNum < 0 || evtNum >= g->cache_len[slot]) {
            g->error = 2;
            continue;
        }
        cplx me  = cmake(w[2 * (int64_t)evtNum], w[2 * (int64_t)evtNum + 1]);
        totalAmp = cadd(totalAmp, cmul(amp, me));
    }
    double ret = totalAmp.re * totalAmp.re + totalAmp.im * totalAmp.im; /* thrust::norm */
    inc5(pc, 1, (int)numResonances * 2, 2, num_obs, 1);
    for(unsigned i = 0; i < numResonances; i++)
        inc(g, pc);
    double eff = call_function(g, evt, pc);
    ret *= eff;
    return ret;
}

/* ------------------------------------------------------- time-dependent Dalitz */

/* device_truth_resolution, src/PDFs/physics/TruthResolution.cu:6-27 */
static double res_truth(double coshterm, double costerm, double sinhterm, double sinterm, double tau, double dtime,
                        double xmixing, double ymixing) {
    double ret = 0;
    dtime /= tau;
    ret += coshterm * cosh(ymixing * dtime);
    ret += costerm * cos(xmixing * dtime);
    ret -= 2 * sinhterm * sinh(ymixing * dtime);
    ret -= 2 * sinterm * sin(xmixing * dtime);
    ret *= exp(-dtime);
    return ret;
}

/* gaussian(), src/PDFs/physics/ThreeGaussResolution.cu:11-61 */
static void tg_gaussian(double *P1, double *P2, double *P3, double *P4, double tau, double adjTime, double xmixing,
                        double ymixing, double adjSigma, double selbias) {
    const double SQRTPIo2  = 1.0 / M_2_SQRTPI;
    const double SQRT1o2PI = sqrt(0.5 * M_1_PI);
    const double R1o6      = 1.0 / 6.0;
    double _1oSqrtA  = adjSigma * M_SQRT2;
    double _1oSigma  = 1 / adjSigma;
    double _1o2SqrtA = 0.5 * _1oSqrtA;
    double _1oSigma2 = _1oSigma * _1oSigma;
    double _NormG    = SQRT1o2PI * _1oSigma;
    double _C        = 0.5 * adjTime * adjTime * _1oSigma2;
    double _Bgn      = -adjTime * _1oSigma2;
    double _Gamma    = 1 / tau;
    double _B        = _Gamma + _Bgn;
    double _u0       = _1o2SqrtA * (_B + selbias);
    double _u02      = _u0 * _u0;
    double _F        = _1oSqrtA * exp(-_C + _u02);
    double _Ig0      = SQRTPIo2 * erfc(_u0);
    double _Ig1      = 0.5 * exp(-_u02);
    double _Ig2      = _Ig1 * _u0 + 0.5 * _Ig0;
    double _Ig3      = _Ig1 * (_u02 + 1);
    double _R        = xmixing * _Gamma * _1oSqrtA;
    double _R2       = _R * _R;
    double _It0      = _F * _Ig0;
    double _It1      = _F * _R * (_Ig1 - _u0 * _Ig0);
    double _It2      = _F * _R2 * (_Ig2 - _u0 * _Ig1 * 2 + _u02 * _Ig0);
    double _It3      = _F * _R2 * _R * (_Ig3 - _u0 * _Ig2 * 3 + _u02 * _Ig1 * 3 - _u0 * _u02 * _Ig0);
    *P2              = _NormG * (_It0 - 0.5 * _It2);
    *P4              = _NormG * (_It1 - R1o6 * _It3);
    double _u0py = _1o2SqrtA * (_B - ymixing * _Gamma + selbias);
    double _u0my = _1o2SqrtA * (_B + ymixing * _Gamma + selbias);
    double _Fpy  = _1oSqrtA * exp(-_C + _u0py * _u0py);
    double _Fmy  = _1oSqrtA * exp(-_C + _u0my * _u0my);
    double _Ipy  = _Fpy * SQRTPIo2 * erfc(_u0py);
    double _Imy  = _Fmy * SQRTPIo2 * erfc(_u0my);
    *P1          = _NormG * 0.5 * (_Ipy + _Imy);
    *P3          = _NormG * 0.5 * (_Ipy - _Imy);
}

/* device_threegauss_resolution, src/PDFs/physics/ThreeGaussResolution.cu:63-127 */
static double res_threegauss(const gfo_prog *g, const pcur *pc, double coshterm, double costerm, double sinhterm,
                             double sinterm, double tau, double dtime, double xmixing, double ymixing,
                             double sigma) {
    double coreFraction    = P(g, pc, 0);
    double tailFraction    = P(g, pc, 1);
    double outlFraction    = 1 - coreFraction - tailFraction;
    double coreBias        = P(g, pc, 2);
    double coreScaleFactor = P(g, pc, 3);
    double tailBias        = P(g, pc, 4);
    double tailScaleFactor = P(g, pc, 5);
    double outlBias        = P(g, pc, 6);
    double outlScaleFactor = P(g, pc, 7);
    double selbias         = P(g, pc, 8);
    double cbo = P(g, pc, 9), cws = P(g, pc, 10), cbo2 = P(g, pc, 11), cws2 = P(g, pc, 12);
    coreBias += cbo;
    tailBias += cbo2;
    outlBias += cbo2;
    coreScaleFactor *= cws;
    tailScaleFactor *= cws2;
    outlScaleFactor *= cws2;
    double cp1 = 0, cp2 = 0, cp3 = 0, cp4 = 0, tp1 = 0, tp2 = 0, tp3 = 0, tp4 = 0, op1 = 0, op2 = 0, op3 = 0, op4 = 0;
    tg_gaussian(&cp1, &cp2, &cp3, &cp4, tau, dtime - coreBias * sigma, xmixing, ymixing, coreScaleFactor * sigma,
                selbias);
    tg_gaussian(&tp1, &tp2, &tp3, &tp4, tau, dtime - tailBias * sigma, xmixing, ymixing, tailScaleFactor * sigma,
                selbias);
    tg_gaussian(&op1, &op2, &op3, &op4, tau, dtime - outlBias * sigma, xmixing, ymixing, outlScaleFactor * sigma,
                selbias);
    double _P1 = coreFraction * cp1 + tailFraction * tp1 + outlFraction * op1;
    double _P2 = coreFraction * cp2 + tailFraction * tp2 + outlFraction * op2;
    double _P3 = coreFraction * cp3 + tailFraction * tp3 + outlFraction * op3;
    double _P4 = coreFraction * cp4 + tailFraction * tp4 + outlFraction * op4;
    double ret = 0;
    ret += coshterm * _P1;
    ret += costerm * _P2;
    ret -= 2 * sinhterm * _P3;
    ret -= 2 * sinterm * _P4;
    return ret;
}

static double call_resolution(gfo_prog *g, const pcur *pc, double t1, double t2, double abr, double abi, double tau,
                              double dtime, double x, double y, double sigma) {
    switch(g->funcs[pc->funcIdx]) {
    case GF_FN_TRUTH_RES: return res_truth(t1, t2, abr, abi, tau, dtime, x, y);
    case GF_FN_THREEGAUSS_RES: return res_threegauss(g, pc, t1, t2, abr, abi, tau, dtime, x, y, sigma);
    default: g->error = 5; return 0;
    }
}

/* device_Tddp, src/PDFs/physics/Amp3Body_TD.cu:41-257: params {tau, x, y, dx, dy,
 * (re_i, im_i)...}, constants {nres, cacheToUse, mistag flag, end of efficiency},
 * observables {dtime, sigma_t, m12, m13, event number[, mistag[, charm tag]]};
 * components: the resonances, the resolution, the efficiency. */
static double f_tddp(gfo_prog *g, const double *evt, pcur *pc) {
    int num_parameters  = NP(g, pc);
    int num_constants   = NC(g, pc);
    int num_observables = NO(g, pc);
    int id_m12 = (int)O(g, pc, 2), id_m13 = (int)O(g, pc, 3), id_num = (int)O(g, pc, 4);
    int id_mis = 0, id_tag = 0;
    if(num_observables > 5)
        id_mis = (int)O(g, pc, 5);
    double m12 = evt[id_m12], m13 = evt[id_m13];
    unsigned numResonances = (unsigned)C(g, pc, 0);
    if(!gfo_in_dalitz(m12, m13, g->motherMass, g->daug1Mass, g->daug2Mass, g->daug3Mass)) {
        unsigned endEfficiencyFunc = (unsigned)C(g, pc, 3);
        inc5(pc, 1, num_parameters, num_constants, num_observables, 1);
        for(unsigned i = 0; i < numResonances; i++)
            inc(g, pc);
        inc(g, pc);
        while(pc->funcIdx < (int)endEfficiencyFunc)
            inc(g, pc);
        return 0;
    }
    int evtNum = (int)floor(0.5 + evt[id_num]);
    cplx sumWavesA = cmake(0, 0), sumWavesB = cmake(0, 0);
    unsigned cacheToUse = (unsigned)C(g, pc, 1);
    for(unsigned i = 0; i < numResonances; ++i) {
        cplx amp      = cmake(P(g, pc, (int)i * 2 + 5), P(g, pc, (int)i * 2 + 6));
        unsigned slot = i + 16 * cacheToUse;
        const double *wa = slot < GFO_MAX_CACHE ? g->cache[slot] : NULL;
        const double *wb = slot < GFO_MAX_CACHE ? g->cache_b[slot] : NULL;
        if(!wa || !wb || evtNum < 0 || evtNum >= g->cache_len[slot]) {
            g->error = 2;
            continue;
        }
        cplx ai   = cmake(wa[2 * (int64_t)evtNum], wa[2 * (int64_t)evtNum + 1]);
        cplx bi   = cmake(wb[2 * (int64_t)evtNum], wb[2 * (int64_t)evtNum + 1]);
        sumWavesA = cadd(sumWavesA, cmul(ai, amp));
        sumWavesB = cadd(sumWavesB, cmul(bi, amp));
    }
    int id_time = (int)O(g, pc, 0), id_sigma = (int)O(g, pc, 1);
    if(num_observables > 6)
        id_tag = (int)O(g, pc, 6);
    double _tau = P(g, pc, 0), _xmixing0 = P(g, pc, 1), _ymixing0 = P(g, pc, 2), _deltax = P(g, pc, 3),
           _deltay = P(g, pc, 4);
    double _xmixing, _ymixing;
    int _charmtag = 0;
    if(num_observables > 6)
        _charmtag = (int)evt[id_tag];
    if(_charmtag == 1) {
        _xmixing = _xmixing0 + _deltax;
        _ymixing = _ymixing0 + _deltay;
    } else if(_charmtag == -1) {
        _xmixing = _xmixing0 - _deltax;
        _ymixing = _ymixing0 - _deltay;
    } else {
        _xmixing = _xmixing0;
        _ymixing = _ymixing0;
    }
    double _time = evt[id_time], _sigma = evt[id_sigma];
    double nA = sumWavesA.re * sumWavesA.re + sumWavesA.im * sumWavesA.im;
    double nB = sumWavesB.re * sumWavesB.re + sumWavesB.im * sumWavesB.im;
    double term1 = nA + nB;
    double term2 = nA - nB;
    sumWavesA    = cmul(sumWavesA, cconj(sumWavesB));
    inc5(pc, 1, num_parameters, num_constants, num_observables, 1);
    for(unsigned i = 0; i < numResonances; i++)
        inc(g, pc);
    double ret = call_resolution(g, pc, term1, term2, sumWavesA.re, sumWavesA.im, _tau, _time, _xmixing, _ymixing,
                                 _sigma);
    double mistag = 0;
    if(num_observables > 5)
        mistag = evt[id_mis];
    const double xfix = 0.0039, yfix = 0.0065, taufix = 0.4101;
    if(mistag > 0) {
        ret *= (1 - 2 * mistag);
        ret += mistag * call_resolution(g, pc, term1, -term2, sumWavesA.re, -sumWavesA.im, taufix, _time, xfix, yfix,
                                        _sigma);
        ret += mistag * call_resolution(g, pc, term1, term2, sumWavesA.re, sumWavesA.im, taufix, _time, xfix, yfix,
                                        _sigma);
    }
    inc(g, pc);
    double eff = call_function(g, evt, pc);
    ret *= eff;
    return ret;
}

/* callFunction, src/PDFs/GooPdf.cu:398-400: dispatch on the function table entry */
static double call_function(gfo_prog *g, const double *evt, pcur *pc) {
    if(pc->funcIdx < 0 || pc->funcIdx >= g->nfunc) {
        g->error = 3;
        return 0;
    }
    switch(g->funcs[pc->funcIdx]) {
    case GF_FN_GAUSSIAN: return f_gaussian(g, evt, pc);
    case GF_FN_EXP: return f_exp(g, evt, pc);
    case GF_FN_EXP_OFFSET: return f_exp_offset(g, evt, pc);
    case GF_FN_EXP_POLY: return f_exp_poly(g, evt, pc);
    case GF_FN_EXP_POLY_OFFSET: return f_exp_poly_offset(g, evt, pc);
    case GF_FN_EXPGAUS: return f_expgaus(g, evt, pc);
    case GF_FN_POLYNOMIAL: return f_polynomial(g, evt, pc);
    case GF_FN_OFFSET_POLYNOMIAL: return f_offset_polynomial(g, evt, pc);
    case GF_FN_MULTI_POLYNOMIAL: return f_multi_polynomial(g, evt, pc);
    case GF_FN_ADD: return f_add(g, evt, pc);
    case GF_FN_ADD_EXT: return f_add_ext(g, evt, pc);
    case GF_FN_PROD: return f_prod(g, evt, pc);
    case GF_FN_DALITZ: return f_dalitz(g, evt, pc);
    case GF_FN_TDDP: return f_tddp(g, evt, pc);
    default: g->error = 4; return 0;
    }
}

/* ------------------------------------------------------------------- metrics */

/* src/PDFs/MetricPointer.cu:8-58; c_totalEvents is the number of entries */
double gfo_metric(int metric, double rawPdf, const double *evtVal, double norm, double totalEvents) {
    switch(metric) {
    case GF_METRIC_EVAL: return rawPdf;
    case GF_METRIC_NLL: rawPdf *= norm; return rawPdf > 0.0 ? -log(rawPdf) : 0.0;
    case GF_METRIC_PROB: return rawPdf * norm;
    case GF_METRIC_BINAVG: {
        rawPdf *= norm;
        rawPdf *= evtVal[1];
        if(rawPdf > 0) {
            double expEvents = totalEvents * rawPdf;
            return (expEvents - evtVal[0] * log(expEvents));
        }
        return 0;
    }
    case GF_METRIC_BINWITHERROR:
        rawPdf -= evtVal[0];
        rawPdf /= evtVal[1];
        rawPdf *= rawPdf;
        return rawPdf;
    case GF_METRIC_CHISQ: {
        rawPdf *= norm;
        rawPdf *= evtVal[1];
        double d = rawPdf * totalEvents - evtVal[0];
        return d * d / (evtVal[0] > 1 ? evtVal[0] : 1);
    }
    default: return NAN;
    }
}

/* ------------------------------------------------------------ program object */

gfo_prog *gfo_prog_new(const int32_t *funcs, int nfunc, const double *params, int np, const double *consts, int nc,
                       const double *obs, int no, const double *norms, int nn, const double *decay6) {
    gfo_prog *g = (gfo_prog *)calloc(1, sizeof(gfo_prog));
    g->funcs    = (int *)malloc(sizeof(int) * (nfunc > 0 ? nfunc : 1));
    for(int i = 0; i < nfunc; i++)
        g->funcs[i] = funcs[i];
    g->nfunc  = nfunc;
    g->params = (double *)malloc(sizeof(double) * (np > 0 ? np : 1));
    memcpy(g->params, params, sizeof(double) * np);
    g->np     = np;
    g->consts = (double *)malloc(sizeof(double) * (nc > 0 ? nc : 1));
    memcpy(g->consts, consts, sizeof(double) * nc);
    g->nc  = nc;
    g->obs = (double *)malloc(sizeof(double) * (no > 0 ? no : 1));
    memcpy(g->obs, obs, sizeof(double) * no);
    g->no    = no;
    g->norms = (double *)malloc(sizeof(double) * (nn > 0 ? nn : 1));
    memcpy(g->norms, norms, sizeof(double) * nn);
    g->nn = nn;
    if(decay6) {
        g->motherMass          = decay6[0];
        g->daug1Mass           = decay6[1];
        g->daug2Mass           = decay6[2];
        g->daug3Mass           = decay6[3];
        g->meson_radius        = decay6[4];
        g->mother_meson_radius = decay6[5];
    }
    return g;
}

void gfo_prog_free(gfo_prog *g) {
    if(!g)
        return;
    free(g->funcs);
    free(g->params);
    free(g->consts);
    free(g->obs);
    free(g->norms);
    free(g);
}

void gfo_prog_set_params(gfo_prog *g, const double *p, int np) {
    if(np == g->np)
        memcpy(g->params, p, sizeof(double) * np);
}
void gfo_prog_set_norms(gfo_prog *g, const double *n, int nn) {
    if(nn == g->nn)
        memcpy(g->norms, n, sizeof(double) * nn);
}
void gfo_prog_set_decay(gfo_prog *g, const double *decay6) {
    g->motherMass          = decay6[0];
    g->daug1Mass           = decay6[1];
    g->daug2Mass           = decay6[2];
    g->daug3Mass           = decay6[3];
    g->meson_radius        = decay6[4];
    g->mother_meson_radius = decay6[5];
}
/* cResonances[slot] = re/im array of len entries (Amp3Body::setDataSize, Amp3Body.cu:222-233) */
void gfo_prog_set_cache(gfo_prog *g, int slot, const double *re_im, int64_t len) {
    if(slot >= 0 && slot < GFO_MAX_CACHE) {
        g->cache[slot]     = re_im;
        g->cache_len[slot] = len;
    }
}
/* second wave of an Amp3Body_TD cache slot: B(m13, m12) beside A(m12, m13) */
void gfo_prog_set_cache_b(gfo_prog *g, int slot, const double *re_im) {
    if(slot >= 0 && slot < GFO_MAX_CACHE)
        g->cache_b[slot] = re_im;
}
int gfo_prog_error(const gfo_prog *g) { return g->error; }
void gfo_prog_set_total_events(gfo_prog *g, double n) { g->total_events = n; }

/* callFunction from the root (function 0) */
double gfo_eval_event(gfo_prog *g, const double *evt) {
    pcur pc = {0, 0, 0, 0, 0};
    return call_function(g, evt, &pc);
}

/* callFunction from function index func_idx, walking the cursor there the slow way
 * (MetricTaker.cu:68-70) */
double gfo_eval_event_at(gfo_prog *g, int func_idx, const double *evt) {
    pcur pc = {0, 0, 0, 0, 0};
    for(int i = 0; i < func_idx; i++)
        inc(g, &pc);
    return call_function(g, evt, &pc);
}

/* Number of OMP chunks thrust would use: omp_get_num_procs() (default_decomposition.inl:50-52) */
int gfo_num_procs(void) {
#ifdef _OPENMP
    return omp_get_num_procs();
#else
    return 1;
#endif
}

/* ------------------------------------------------ unbinned metric (MetricTaker) */

/* MetricTaker::operator()(tuple<int, fptype*, int>), MetricTaker.cu:17-44 */
static double metric_taker(gfo_prog *g, const double *base, int64_t i, int eventSize, int metric) {
    pcur pc             = {0, 0, 0, 0, 0};
    const double *evt   = base + i * (int64_t)abs(eventSize);
    double norm         = N(g, &pc, 0);
    double ret          = call_function(g, evt, &pc);
    return gfo_metric(metric, ret, evt + (abs(eventSize) - 2), norm, g->total_events);
}

void gfo_eval_metric(gfo_prog *g, const double *aos, int64_t n, int evtSize, int metric, double *out) {
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for(int64_t i = 0; i < n; i++)
        out[i] = metric_taker(g, aos, i, evtSize, metric);
}

/* uniform_decomposition(n, 1, P): first (n % P) intervals get one extra element
 * (system/detail/internal/decompose.h:67-97) */
static void chunk_bounds(int64_t n, int P, int c, int64_t *b, int64_t *e) {
    int64_t intervals = n;
    int64_t small = 1, large = 0, threshold = 0;
    if(intervals > P) {
        small     = n / P;
        large     = small + 1;
        threshold = n % P;
        intervals = P;
    }
    (void)intervals;
    if(c < threshold) {
        *b = large * c;
        *e = *b + large;
    } else {
        *b = large * threshold + small * (c - threshold);
        *e = (*b + small < n) ? *b + small : n;
    }
}

int gfo_num_chunks(int64_t n, int P) { return (int)(n > P ? P : n); }

/* thrust::transform_reduce on the OMP backend (reduce.inl:51-66): chunk sums from
 * each chunk's first element, then init(0.0) + s_1 + ... + s_P in order. */
double gfo_reduce_metric(gfo_prog *g, const double *aos, int64_t n, int evtSize, int metric, int P) {
    if(n <= 0)
        return 0.0;
    int nch        = gfo_num_chunks(n, P);
    double *chunks = (double *)malloc(sizeof(double) * nch);
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for(int c = 0; c < nch; c++) {
        int64_t b, e;
        chunk_bounds(n, P, c, &b, &e);
        double s = metric_taker(g, aos, b, evtSize, metric);
        for(int64_t i = b + 1; i < e; i++)
            s = s + metric_taker(g, aos, i, evtSize, metric);
        chunks[c] = s;
    }
    double ret = 0.0;
    for(int c = 0; c < nch; c++)
        ret = ret + chunks[c];
    free(chunks);
    return ret;
}

/* Same chunked order over a precomputed array of terms */
double gfo_chunked_sum(const double *terms, int64_t n, int P) {
    if(n <= 0)
        return 0.0;
    int nch    = gfo_num_chunks(n, P);
    double ret = 0.0;
    for(int c = 0; c < nch; c++) {
        int64_t b, e;
        chunk_bounds(n, P, c, &b, &e);
        double s = terms[b];
        for(int64_t i = b + 1; i < e; i++)
            s = s + terms[i];
        ret = ret + s;
    }
    return ret;
}

/* ---------------------------------------------------- binned grid (MetricTaker) */

/* Value of function func_idx at global bin `bin` of the grid ranges[3k].
 * MetricTaker::operator()(tuple<int,int,fptype*>), MetricTaker.cu:48-96.  For
 * k == 1 this is the reference exactly (slot events[id] holds the bin centre).
 * For k > 1 the reference writes every coordinate into the same slot
 * events[getObservable(0)] and leaves the other slots uninitialised; here each
 * coordinate goes to the slot of its own observable (the midpoint rule). */
/* Centres of global bin binNumber in each of the k grid dimensions (MetricTaker.cu:76-86:
 * first dimension fastest, x = ((u - l) / n) * (localBin + 0.5) + l). */
static void grid_centres(const double *ranges, int k, int64_t binNumber, double *x) {
    for(int i = 0; i < k; ++i) {
        double lowerBound = ranges[3 * i + 0];
        double upperBound = ranges[3 * i + 1];
        int numBins       = (int)floor(ranges[3 * i + 2] + 0.5);
        int localBin      = (int)(binNumber % numBins);
        double c          = upperBound - lowerBound;
        c /= numBins;
        c *= (localBin + 0.5);
        c += lowerBound;
        x[i] = c;
        binNumber /= numBins;
    }
}

static double binned_value(gfo_prog *g, int func_idx, const double *ranges, int k, int64_t binNumber) {
    pcur pc = {0, 0, 0, 0, 0};
    double events[16], x[16];
    memset(events, 0, sizeof(events));
    for(int i = 0; i < func_idx; i++)
        inc(g, &pc);
    grid_centres(ranges, k, binNumber, x);
    for(int i = 0; i < k; ++i) {
        int id = (int)O(g, &pc, k == 1 ? 0 : i);
        if(id >= 0 && id < 16)
            events[id] = x[i];
    }
    return call_function(g, events, &pc);
}

/* The cell centres of a k-dimensional grid: out[d * total + b] = coordinate d of bin b. */
void gfo_grid_cells(const double *ranges, int k, double *out) {
    int64_t total = 1;
    for(int i = 0; i < k; i++)
        total *= (int64_t)floor(ranges[3 * i + 2] + 0.5);
    for(int64_t b = 0; b < total; b++) {
        double x[16];
        grid_centres(ranges, k, b, x);
        for(int d = 0; d < k; d++)
            out[(int64_t)d * total + b] = x[d];
    }
}

double gfo_reduce_bins(gfo_prog *g, int func_idx, const double *ranges, int k, int P) {
    int64_t total = 1;
    for(int i = 0; i < k; i++)
        total *= (int64_t)floor(ranges[3 * i + 2] + 0.5);
    if(total <= 0)
        return 0.0;
    int nch        = gfo_num_chunks(total, P);
    double *chunks = (double *)malloc(sizeof(double) * nch);
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for(int c = 0; c < nch; c++) {
        int64_t b, e;
        chunk_bounds(total, P, c, &b, &e);
        double s = binned_value(g, func_idx, ranges, k, b);
        for(int64_t i = b + 1; i < e; i++)
            s = s + binned_value(g, func_idx, ranges, k, i);
        chunks[c] = s;
    }
    double ret = 0.0;
    for(int c = 0; c < nch; c++)
        ret = ret + chunks[c];
    free(chunks);
    return ret;
}

void gfo_eval_bins(gfo_prog *g, int func_idx, const double *ranges, int k, double *out) {
    int64_t total = 1;
    for(int i = 0; i < k; i++)
        total *= (int64_t)floor(ranges[3 * i + 2] + 0.5);
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for(int64_t i = 0; i < total; i++)
        out[i] = binned_value(g, func_idx, ranges, k, i);
}

/* ----------------------------------------------------------- Dalitz caching */

/* Lineshape of the resonance with function index res_func at (m12, m13):
 * 0 outside the Dalitz region.  SpecialResonanceCalculator::operator(),
 * detail/SpecialResonanceCalculator.cu:12-46. */
void gfo_lineshape(gfo_prog *g, int res_func, double m12, double m13, double *out_re_im) {
    cplx ret = cmake(0, 0);
    if(gfo_in_dalitz(m12, m13, g->motherMass, g->daug1Mass, g->daug2Mass, g->daug3Mass)) {
        double m23 = g->motherMass * g->motherMass + g->daug1Mass * g->daug1Mass + g->daug2Mass * g->daug2Mass
                     + g->daug3Mass * g->daug3Mass - m12 - m13;
        pcur pc = {0, 0, 0, 0, 0};
        while(pc.funcIdx < res_func)
            inc(g, &pc);
        ret = resonance_amplitude(g, m12, m13, m23, &pc);
    }
    out_re_im[0] = ret.re;
    out_re_im[1] = ret.im;
}

/* Wave cache of one resonance over event rows offset..offset+n-1 (Amp3Body.cu:292-310):
 * the Dalitz node's observables give the m12/m13 slots of each AoS event. */
void gfo_dalitz_wave(gfo_prog *g, int dalitz_func, int res_func, const double *aos, int64_t n, int evtSize,
                     int64_t offset, double *out_re_im) {
    pcur pc = {0, 0, 0, 0, 0};
    while(pc.funcIdx < dalitz_func)
        inc(g, &pc);
    int id_m12 = (int)O(g, &pc, 0);
    int id_m13 = (int)O(g, &pc, 1);
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for(int64_t k = 0; k < n; k++) {
        const double *evt = aos + (offset + k) * evtSize;
        gfo_lineshape(g, res_func, evt[id_m12], evt[id_m13], out_re_im + 2 * k);
    }
}

/* Move the cursor past the whole subtree rooted at pc->funcIdx without
 * evaluating it; the child counts are those the device functions walk
 * (AddPdf.cu:16-79, ProdPdf.cu:9-26, Amp3Body.cu:60-122). */
static void skip_subtree(gfo_prog *g, pcur *pc) {
    int f = g->funcs[pc->funcIdx];
    int nchild = 0;
    if(f == GF_FN_ADD)
        nchild = NP(g, pc) + 1;
    else if(f == GF_FN_ADD_EXT)
        nchild = NP(g, pc);
    else if(f == GF_FN_PROD)
        nchild = (int)C(g, pc, 0);
    else if(f == GF_FN_DALITZ)
        nchild = (int)C(g, pc, 0) + 1;
    else if(f == GF_FN_TDDP)
        nchild = (int)C(g, pc, 0) + 2; /* resonances, resolution, efficiency */
    inc(g, pc);
    for(int i = 0; i < nchild; i++)
        skip_subtree(g, pc);
}

/* Function indices of the resonances of a Dalitz node and of its efficiency:
 * components follow the node in depth-first order (Amp3Body.cu:167-173), the
 * efficiency index is the last function of the subtree (Amp3Body::populateArrays,
 * Amp3Body.cu:198-202). Returns the number of resonances. */
int gfo_dalitz_layout(gfo_prog *g, int dalitz_func, int *res_funcs, int *eff_func) {
    pcur pc = {0, 0, 0, 0, 0};
    while(pc.funcIdx < dalitz_func)
        inc(g, &pc);
    int nres = (int)C(g, &pc, 0);
    inc(g, &pc);
    for(int i = 0; i < nres; i++) {
        res_funcs[i] = pc.funcIdx;
        inc(g, &pc);
    }
    skip_subtree(g, &pc);
    *eff_func = pc.funcIdx - 1;
    return nres;
}

/* Centre of Dalitz integration cell `bin` (SpecialResonanceIntegrator.cu:59-76): m12 bin
 * bin % n12, m13 bin bin / n12; the time-integrated integrator takes the m13 bin count
 * from the m12 entry (hazard A3, :72), the time-dependent one its own
 * (SpecialDalitzIntegrator.cu:75-82).  The cell index is an int there. */
static void dalitz_cell(const double *grid6, int td, int64_t bin, double *m12, double *m13) {
    int globalBinNumber  = (int)bin;
    double lowerBoundM12 = grid6[0];
    double upperBoundM12 = grid6[1];
    int numBinsM12       = (int)floor(grid6[2] + 0.5);
    int binNumberM12     = globalBinNumber % numBinsM12;
    double binCenterM12  = upperBoundM12 - lowerBoundM12;
    binCenterM12 /= numBinsM12;
    binCenterM12 *= (binNumberM12 + 0.5);
    binCenterM12 += lowerBoundM12;
    globalBinNumber /= numBinsM12;
    double lowerBoundM13 = grid6[3];
    double upperBoundM13 = grid6[4];
    int numBinsM13       = (int)floor(grid6[td ? 5 : 2] + 0.5);
    double binCenterM13  = upperBoundM13 - lowerBoundM13;
    binCenterM13 /= numBinsM13;
    binCenterM13 *= (globalBinNumber + 0.5);
    binCenterM13 += lowerBoundM13;
    *m12 = binCenterM12;
    *m13 = binCenterM13;
}

void gfo_dalitz_grid_cells(const double *grid6, int td, double *m12, double *m13) {
    int64_t total = (int64_t)floor(grid6[2] + 0.5) * (int64_t)floor(grid6[5] + 0.5);
    for(int64_t b = 0; b < total; b++)
        dalitz_cell(grid6, td, b, &m12[b], &m13[b]);
}

/* Grid integrals I_ij for all pairs (Amp3Body.cu:312-330 with
 * SpecialResonanceIntegrator.cu:52-121): the bin centres, the m13 bin count read
 * from the m12 entry (SpecialResonanceIntegrator.cu:72), the fake event for the
 * efficiency, and each pair's OMP-chunked complex sum.  eff_func is the
 * efficiencyFunction index.  out: nres*nres complex, row-major. */
void gfo_dalitz_integrals(gfo_prog *g, int dalitz_func, const double *grid6, int eff_func, int P,
                          double *out_re_im) {
    int res_funcs[64];
    int dummy_eff;
    int nres = gfo_dalitz_layout(g, dalitz_func, res_funcs, &dummy_eff);
    pcur dpc = {0, 0, 0, 0, 0};
    while(dpc.funcIdx < dalitz_func)
        inc(g, &dpc);
    int id_m12 = (int)O(g, &dpc, 0);
    int id_m13 = (int)O(g, &dpc, 1);
    int n12    = (int)floor(grid6[2] + 0.5);
    int n13    = (int)floor(grid6[5] + 0.5);
    int64_t totalBins = (int64_t)n12 * (int64_t)n13;
    int nch           = gfo_num_chunks(totalBins, P);
    int npair         = nres * nres;
    cplx *chunk       = (cplx *)calloc((size_t)nch * npair, sizeof(cplx));
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for(int c = 0; c < nch; c++) {
        int64_t b, e;
        chunk_bounds(totalBins, P, c, &b, &e);
        cplx B[64];
        cplx *acc = chunk + (size_t)c * npair;
        for(int64_t bin = b; bin < e; bin++) {
            double binCenterM12, binCenterM13;
            dalitz_cell(grid6, 0, bin, &binCenterM12, &binCenterM13);

            int inside = gfo_in_dalitz(binCenterM12, binCenterM13, g->motherMass, g->daug1Mass, g->daug2Mass,
                                       g->daug3Mass);
            double m23 = g->motherMass * g->motherMass + g->daug1Mass * g->daug1Mass + g->daug2Mass * g->daug2Mass
                         + g->daug3Mass * g->daug3Mass - binCenterM12 - binCenterM13;
            for(int r = 0; r < nres; r++) {
                if(inside) {
                    pcur rp = dpc;
                    while(rp.funcIdx < res_funcs[r])
                        inc(g, &rp);
                    B[r] = resonance_amplitude(g, binCenterM12, binCenterM13, m23, &rp);
                } else {
                    B[r] = cmake(0, 0);
                }
            }
            double events[16];
            memset(events, 0, sizeof(events));
            events[0] = 2;
            if(id_m12 >= 0 && id_m12 < 16)
                events[id_m12] = binCenterM12;
            if(id_m13 >= 0 && id_m13 < 16)
                events[id_m13] = binCenterM13;
            pcur epc = dpc;
            while(epc.funcIdx < eff_func)
                inc(g, &epc);
            double eff = call_function(g, events, &epc);
            for(int i = 0; i < nres; i++) {
                for(int j = 0; j < nres; j++) {
                    cplx v = inside ? cmul(B[i], cconj(B[j])) : cmake(0, 0);
                    v      = cscale(v, eff);
                    if(bin == b)
                        acc[i * nres + j] = v;
                    else
                        acc[i * nres + j] = cadd(acc[i * nres + j], v);
                }
            }
        }
    }
    (void)n13;
    for(int p = 0; p < npair; p++) {
        cplx s = cmake(0, 0);
        for(int c = 0; c < nch; c++)
            s = cadd(s, chunk[(size_t)c * npair + p]);
        out_re_im[2 * p]     = s.re;
        out_re_im[2 * p + 1] = s.im;
    }
    free(chunk);
}

/* Amp3Body::normalize host assembly (Amp3Body.cu:332-358): Re sum_ij c_i conj(c_j) I_ij
 * times the two bin sizes.  coup: nres complex couplings (re, im). */
double gfo_dalitz_norm_from_integrals(int nres, const double *coup, const double *I, double binSize12,
                                      double binSize13) {
    cplx sumIntegral = cmake(0, 0);
    for(int i = 0; i < nres; ++i) {
        cplx ai = cmake(coup[2 * i], coup[2 * i + 1]);
        for(int j = 0; j < nres; ++j) {
            cplx aj  = cmake(coup[2 * j], -coup[2 * j + 1]);
            cplx Iij = cmake(I[2 * (i * nres + j)], I[2 * (i * nres + j) + 1]);
            sumIntegral = cadd(sumIntegral, cmul(cmul(ai, aj), Iij));
        }
    }
    double ret           = sumIntegral.re;
    double binSizeFactor = 1;
    binSizeFactor *= binSize12;
    binSizeFactor *= binSize13;
    ret *= binSizeFactor;
    return ret;
}

/* ------------------------------------------------------- Amp3Body_TD host side */

/* Layout of an Amp3Body_TD node: its resonances, the resolution function, the
 * root of the efficiency subtree (efficiencyFunction, Amp3Body_TD.cu:518)
 * and the function index after that subtree (the constant the device function
 * walks to, Amp3Body_TD.cu:523-527).  Returns the number of resonances. */
int gfo_tddp_layout(gfo_prog *g, int tddp_func, int *res_funcs, int *res_func, int *eff_func, int *end_func) {
    pcur pc = {0, 0, 0, 0, 0};
    while(pc.funcIdx < tddp_func)
        inc(g, &pc);
    int nres = (int)C(g, &pc, 0);
    inc(g, &pc);
    for(int i = 0; i < nres; i++) {
        res_funcs[i] = pc.funcIdx;
        inc(g, &pc);
    }
    *res_func = pc.funcIdx;
    skip_subtree(g, &pc);
    *eff_func = pc.funcIdx;
    skip_subtree(g, &pc);
    *end_func = pc.funcIdx;
    return nres;
}

/* (A, B) waves of one resonance at one point: A = BW(m12, m13), B = BW(m13, m12)
 * with m23 taken once from (m12, m13); both 0 outside the Dalitz region.
 * SpecialWaveCalculator::operator(), detail/SpecialWaveCalculator.cu:13-63. */
static void tddp_ab(gfo_prog *g, const pcur *rpc, double m12, double m13, cplx *a, cplx *b) {
    *a = cmake(0, 0);
    *b = cmake(0, 0);
    if(!gfo_in_dalitz(m12, m13, g->motherMass, g->daug1Mass, g->daug2Mass, g->daug3Mass))
        return;
    double m23 = g->motherMass * g->motherMass + g->daug1Mass * g->daug1Mass + g->daug2Mass * g->daug2Mass
                 + g->daug3Mass * g->daug3Mass - m12 - m13;
    pcur t = *rpc;
    *a     = resonance_amplitude(g, m12, m13, m23, &t);
    t      = *rpc;
    *b     = resonance_amplitude(g, m13, m12, m23, &t);
}

/* Wave cache of one resonance of a TD node over event rows offset..offset+n-1:
 * the m12/m13 slots are the node's observables 2 and 3 (Amp3Body_TD.cu:639-661). */
void gfo_tddp_wave(gfo_prog *g, int tddp_func, int res_func, const double *aos, int64_t n, int evtSize,
                   int64_t offset, double *out_a, double *out_b) {
    pcur pc = {0, 0, 0, 0, 0};
    while(pc.funcIdx < tddp_func)
        inc(g, &pc);
    int id_m12 = (int)O(g, &pc, 2);
    int id_m13 = (int)O(g, &pc, 3);
    pcur rpc   = pc;
    while(rpc.funcIdx < res_func)
        inc(g, &rpc);
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for(int64_t k = 0; k < n; k++) {
        const double *evt = aos + (offset + k) * evtSize;
        cplx a, b;
        tddp_ab(g, &rpc, evt[id_m12], evt[id_m13], &a, &b);
        out_a[2 * k]     = a.re;
        out_a[2 * k + 1] = a.im;
        out_b[2 * k]     = b.re;
        out_b[2 * k + 1] = b.im;
    }
}

/* ThreeComplex integrals of every resonance pair over the m12 x m13 midpoint
 * grid (Amp3Body_TD.cu:664-682 with SpecialDalitzIntegrator.cu:9-126): per cell
 * (ai aj*, ai bj*, bi bj*) times the efficiency evaluated on a fake event
 * {2, ..., m12, m13} (other slots 0 here); the m13 bin count is grid6[5].  Each
 * pair's sum is OMP-chunked like the other oracle sums.  out_*: nres*nres
 * complex, row-major. */
void gfo_tddp_integrals(gfo_prog *g, int tddp_func, const double *grid6, int P, double *out_aa, double *out_ab,
                        double *out_bb) {
    int res_funcs[64], res_func, eff_func, end_func;
    int nres = gfo_tddp_layout(g, tddp_func, res_funcs, &res_func, &eff_func, &end_func);
    pcur dpc = {0, 0, 0, 0, 0};
    while(dpc.funcIdx < tddp_func)
        inc(g, &dpc);
    int id_m12 = (int)O(g, &dpc, 2);
    int id_m13 = (int)O(g, &dpc, 3);
    pcur rpc[64];
    for(int r = 0; r < nres; r++) {
        rpc[r] = dpc;
        while(rpc[r].funcIdx < res_funcs[r])
            inc(g, &rpc[r]);
    }
    pcur epc = dpc;
    while(epc.funcIdx < eff_func)
        inc(g, &epc);
    int n12           = (int)floor(grid6[2] + 0.5);
    int n13           = (int)floor(grid6[5] + 0.5);
    int64_t totalBins = (int64_t)n12 * (int64_t)n13;
    int nch           = gfo_num_chunks(totalBins, P);
    int npair         = nres * nres;
    cplx *chunk       = (cplx *)calloc((size_t)nch * npair * 3, sizeof(cplx));
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for(int c = 0; c < nch; c++) {
        int64_t b0, e0;
        chunk_bounds(totalBins, P, c, &b0, &e0);
        cplx A[64], B[64];
        cplx *acc = chunk + (size_t)c * npair * 3;
        for(int64_t bin = b0; bin < e0; bin++) {
            double binCenterM12, binCenterM13;
            dalitz_cell(grid6, 1, bin, &binCenterM12, &binCenterM13);
            for(int r = 0; r < nres; r++)
                tddp_ab(g, &rpc[r], binCenterM12, binCenterM13, &A[r], &B[r]);
            double events[16];
            memset(events, 0, sizeof(events));
            events[0] = 2;
            if(id_m12 >= 0 && id_m12 < 16)
                events[id_m12] = binCenterM12;
            if(id_m13 >= 0 && id_m13 < 16)
                events[id_m13] = binCenterM13;
            pcur e  = epc;
            double eff = call_function(g, events, &e);
            for(int i = 0; i < nres; i++) {
                for(int j = 0; j < nres; j++) {
                    cplx v[3];
                    v[0] = cscale(cmul(A[i], cconj(A[j])), eff);
                    v[1] = cscale(cmul(A[i], cconj(B[j])), eff);
                    v[2] = cscale(cmul(B[i], cconj(B[j])), eff);
                    for(int t = 0; t < 3; t++) {
                        cplx *dst = acc + (size_t)t * npair + i * nres + j;
                        *dst      = (bin == b0) ? v[t] : cadd(*dst, v[t]);
                    }
                }
            }
        }
    }
    double *outs[3] = {out_aa, out_ab, out_bb};
    for(int t = 0; t < 3; t++)
        for(int p = 0; p < npair; p++) {
            cplx s = cmake(0, 0);
            for(int c = 0; c < nch; c++)
                s = cadd(s, chunk[((size_t)c * 3 + t) * npair + p]);
            outs[t][2 * p]     = s.re;
            outs[t][2 * p + 1] = s.im;
        }
    free(chunk);
}

/* TruthResolution::normalization, src/PDFs/physics/TruthResolution.cu:58-71 */
static double norm_truth(double di1, double di2, double di3, double di4, double tau, double xmixing,
                         double ymixing) {
    double timeIntegralOne = tau / (1 - ymixing * ymixing);
    double timeIntegralTwo = tau / (1 + xmixing * xmixing);
    double timeIntegralThr = ymixing * timeIntegralOne;
    double timeIntegralFou = xmixing * timeIntegralTwo;
    double ret             = timeIntegralOne * (di1 + di2);
    ret += timeIntegralTwo * (di1 - di2);
    ret -= 2 * timeIntegralThr * di3;
    ret -= 2 * timeIntegralFou * di4;
    return ret;
}

/* ThreeGaussResolution::normalization, src/PDFs/physics/ThreeGaussResolution.cu:153-182 */
static double norm_threegauss(double di1, double di2, double di3, double di4, double tau, double xmixing,
                              double ymixing, double selBias) {
    double timeIntegralOne
        = (selBias + 1 / tau) / (selBias * selBias + 2 * selBias / tau + (1 - ymixing * ymixing) / (tau * tau));
    double timeIntegralTwo
        = (selBias + 1 / tau) / (selBias * selBias + 2 * selBias / tau + (1 + xmixing * xmixing) / (tau * tau));
    double timeIntegralThr
        = (ymixing / tau) / (selBias * selBias + 2 * selBias / tau + (1 - ymixing * ymixing) / (tau * tau));
    double timeIntegralFou
        = (xmixing / tau) / (selBias * selBias + 2 * selBias / tau + (1 + xmixing * xmixing) / (tau * tau));
    double ret = timeIntegralOne * (di1 + di2);
    ret += timeIntegralTwo * (di1 - di2);
    ret -= 2 * timeIntegralThr * di3;
    ret -= 2 * timeIntegralFou * di4;
    return ret;
}

/* MixingTimeResolution::normalization of a resolution kind (a GF_FN_*_RES code);
 * selBias is the ThreeGauss selection-bias value. */
double gfo_resolution_norm(int kind, double di1, double di2, double di3, double di4, double tau, double x, double y,
                           double selBias) {
    if(kind == GF_FN_THREEGAUSS_RES)
        return norm_threegauss(di1, di2, di3, di4, tau, x, y, selBias);
    return norm_truth(di1, di2, di3, di4, tau, x, y);
}

/* Resolution function value at one decay time (the device functions above):
 * params13 are the ThreeGauss parameters (ignored for the truth resolution). */
double gfo_resolution_eval(int kind, const double *params13, double t1, double t2, double abr, double abi,
                           double tau, double dtime, double x, double y, double sigma) {
    if(kind == GF_FN_THREEGAUSS_RES) {
        gfo_prog g;
        memset(&g, 0, sizeof(g));
        double pblock[14];
        pblock[0] = 13;
        memcpy(pblock + 1, params13, 13 * sizeof(double));
        g.params = pblock;
        pcur pc  = {0, 0, 0, 0, 0};
        return res_threegauss(&g, &pc, t1, t2, abr, abi, tau, dtime, x, y, sigma);
    }
    return res_truth(t1, t2, abr, abi, tau, dtime, x, y);
}

/* Amp3Body_TD::normalize host assembly (Amp3Body_TD.cu:686-751).  coup: nres
 * (re, im); I*: the ThreeComplex integrals; mix5 = {tau, x, y, dx, dy};
 * mistag_on / mistag: whether the mistag observable tests true (its value is
 * in range, include/goofit/Variable.h:103-105) and its value. */
double gfo_tddp_norm_from_integrals(int nres, const double *coup, const double *Iaa, const double *Iab,
                                    const double *Ibb, const double *mix5, int res_kind, double selBias,
                                    double d0fraction, int mistag_on, double mistag, double binSize12,
                                    double binSize13) {
    cplx integralA_2 = cmake(0, 0), integralB_2 = cmake(0, 0), integralABs = cmake(0, 0);
    for(int i = 0; i < nres; ++i) {
        cplx amplitude_i = cmake(coup[2 * i], coup[2 * i + 1]);
        for(int j = 0; j < nres; ++j) {
            cplx amplitude_j = cmake(coup[2 * j], -coup[2 * j + 1]);
            int k            = 2 * (i * nres + j);
            integralA_2 = cadd(integralA_2, cmul(cmul(amplitude_i, amplitude_j), cmake(Iaa[k], Iaa[k + 1])));
            integralABs = cadd(integralABs, cmul(cmul(amplitude_i, amplitude_j), cmake(Iab[k], Iab[k + 1])));
            integralB_2 = cadd(integralB_2, cmul(cmul(amplitude_i, amplitude_j), cmake(Ibb[k], Ibb[k + 1])));
        }
    }
    double one = integralA_2.re, two = integralB_2.re, thr = integralABs.re, fou = integralABs.im;
    double tau = mix5[0], x0 = mix5[1], y0 = mix5[2], dx = mix5[3], dy = mix5[4];
    double ret_D0    = gfo_resolution_norm(res_kind, one, two, thr, fou, tau, x0 + dx, y0 + dy, selBias);
    double ret_D0bar = gfo_resolution_norm(res_kind, one, two, thr, fou, tau, x0 - dx, y0 - dy, selBias);
    double ret       = d0fraction * ret_D0 + (1. - d0fraction) * ret_D0bar;
    const double xfix = 0.0039, yfix = 0.0065, taufix = 0.4101;
    if(mistag_on) {
        ret *= (1 - 2 * mistag);
        ret += mistag * gfo_resolution_norm(res_kind, one, two, thr, fou, taufix, xfix, yfix, selBias);
        ret += mistag * gfo_resolution_norm(res_kind, two, one, thr, -fou, taufix, xfix, yfix, selBias);
    }
    double binSizeFactor = 1;
    binSizeFactor *= binSize12;
    binSizeFactor *= binSize13;
    ret *= binSizeFactor;
    return ret;
}

/* ------------------------------------------------------------ exact sums */

/* Neumaier-compensated sum in long double of an array: a near-exact reference
 * used to check both the chunked oracle sum and the GPU's compensated sum. */
double gfo_exact_sum(const double *x, int64_t n) {
    long double s = 0.0L, c = 0.0L;
    for(int64_t i = 0; i < n; i++) {
        long double t = s + (long double)x[i];
        if(fabsl(s) >= fabsl((long double)x[i]))
            c += (s - t) + (long double)x[i];
        else
            c += ((long double)x[i] - t) + s;
        s = t;
    }
    return (double)(s + c);
}
