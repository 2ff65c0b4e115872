// gf_math.h -- FP64 log for the likelihood metrics, and the NLL of many events as
// one logarithm.
//
// Every event of every likelihood sweep takes one logarithm per parameter point
// (-log of the normalised PDF value), and the multi-point sweeps -- up to 16
// points per event, MIGRAD's derivative rounds and the toy batches -- are bound by
// FP64 issue.  The device library's log spends about 76 FP64 instructions (much of
// it extra-precision arithmetic); this one spends 40 and one division, with an
// error below one unit in the last place over the whole double range (measured:
// 0.9 at most) and the IEEE results at 0, +inf, negative numbers and NaN.  On
// gfx950 it runs 1.5 times the device library's throughput.  It uses only
// correctly rounded operations (fma, frexp, +, *, /, comparisons and selects), so a
// host build of this header (GF_MATH_HD empty under a plain C++ compiler) gives the
// same bits as the device; the accuracy is pinned on the host against long double.
#pragma once

#include <cmath>

#ifdef __HIPCC__
#define GF_MATH_HD __host__ __device__
#else
#define GF_MATH_HD
#endif

namespace gfmath {

// ln 2 = ln2_hi + ln2_lo: ln2_hi is ln 2 rounded to 32 fractional bits (28 significant
// bits), so e * ln2_hi is exact for every binary exponent e of a double; ln2_lo is the
// rest, rounded to double.
constexpr double ln2_hi = 0x1.62e42ffp-1;
constexpr double ln2_lo = -0x1.718432a1b0e26p-35;

// log(x).  x = m 2^e with m in [sqrt(1/2), sqrt(2)) (frexp, then one doubling); with
// f = m - 1 (exact) and s = f / (2 + f), log m = 2 atanh(s) = f - s (f - Q), where
// Q = 2 s^2 (1/3 + s^2/5 + s^4/7 + ...) -- the atanh series, |s| < 0.1716, summed to
// s^20 (the first omitted term is below 2^-57 of log m) -- since 2 s = f - s f.  The
// form keeps f exact and lets the roundings of s enter only through the correction
// s (f - Q), at most a fifth of the result.  For |e| >= 2, log x = e ln2_hi + (e ln2_lo
// + log m), one rounding of a sum that log m enters at a quarter at most.  For |e| = 1
// (x in [sqrt(2)/4, sqrt(2)/2) or [sqrt(2), 2 sqrt(2))), where e ln 2 and log m cancel
// down to half, log x = (e ln2_hi + f) - (s (f - Q) - e ln2_lo): the first sum is exact
// there (e ln2_hi has 28 significant bits and f none below the sum's last place), so
// only the small correction carries rounding error.  log(+inf) = +inf, log(0) = -inf,
// log(x < 0) = NaN, NaN propagates; subnormal inputs are normalised by frexp.
GF_MATH_HD inline double log_d(double x) {
    int e;
    double m      = frexp(x, &e);  // x = m 2^e, 0.5 <= |m| < 1 for finite nonzero x
    const bool lo = m < 0x1.6a09e667f3bcdp-1;  // sqrt(1/2)
    m             = lo ? m + m : m;
    const double ed = (double)(lo ? e - 1 : e);
    const double f  = m - 1.0;
    const double s  = f / (2.0 + f);
    const double z  = s * s;
    // 2/(2n+1) for n = 10 .. 1, each rounded to double
    double q = 0x1.8618618618618p-4;
    q        = fma(q, z, 0x1.af286bca1af28p-4);
    q        = fma(q, z, 0x1.e1e1e1e1e1e1ep-4);
    q        = fma(q, z, 0x1.1111111111111p-3);
    q        = fma(q, z, 0x1.3b13b13b13b14p-3);
    q        = fma(q, z, 0x1.745d1745d1746p-3);
    q        = fma(q, z, 0x1.c71c71c71c71cp-3);
    q        = fma(q, z, 0x1.2492492492492p-2);
    q        = fma(q, z, 0x1.999999999999ap-2);
    q        = fma(q, z, 0x1.5555555555555p-1);
    const double Q    = z * q;
    const double logm = fma(-s, f - Q, f);
    const double far  = fma(ed, ln2_hi, fma(ed, ln2_lo, logm));
    const double near = fma(ed, ln2_hi, f) - fma(s, f - Q, -(ed * ln2_lo));
    double y          = (ed == 1.0 || ed == -1.0) ? near : far;
    y                 = x == 0.0 ? -HUGE_VAL : y;
    y                 = (x < 0.0 || x != x) ? NAN : y;
    y                 = x == HUGE_VAL ? HUGE_VAL : y;
    return y;
}

// The sum of -log(v) over many values as one logarithm: sum -log(v_k) over the v_k > 0
// (the others add 0, as the NLL metric gives them) equals -log(prod v_k).  The product
// is kept as a mantissa m and a binary exponent e (an integer, held in a double: exact
// below 2^53): each factor's own mantissa (frexp, exact) multiplies m (one rounding),
// the result is renormalised into [0.5, 1) (frexp, exact) and the exponents are added,
// so the product neither overflows nor underflows, and each factor adds one rounding of
// relative size 2^-53 to it -- an absolute error in the sum of logarithms below that of
// taking each logarithm apart.  A factor v <= 0 or NaN becomes 1.0 = 0.5 * 2^1, whose
// halving is exact and undone by the renormalisation; v = +inf makes the product +inf
// (the sum -inf), as the per-value form gives.  Start from (m, e) = (1, 0).
GF_MATH_HD inline void nll_add(double &m, double &e, double v) {
    int ev, em;
    const double a = frexp(v > 0.0 ? v : 1.0, &ev);
    m              = frexp(m * a, &em);
    e += (double)(ev + em);
}
// The same product with the exponent in an integer register (no conversion per factor):
// m is the double form's m bit for bit and e its exponent's value, for any sequence of
// factors whose exponent sum stays within int range (+-2^31: over 2^20 factors at the
// extreme exponents).
GF_MATH_HD inline void nll_add(double &m, int &e, double v) {
    int ev, em;
    const double a = frexp(v > 0.0 ? v : 1.0, &ev);
    m              = frexp(m * a, &em);
    e += ev + em;
}
// sum of -log(v_k) = -(log m + e ln 2)
GF_MATH_HD inline double nll_total(double m, double e) { return -fma(e, ln2_hi, fma(e, ln2_lo, log_d(m))); }

}  // namespace gfmath
