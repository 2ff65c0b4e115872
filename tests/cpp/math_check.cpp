// Accuracy of gfmath::log_d (goofit_amd/csrc/gf_math.h, built for the host with the
// same operations as on the device) against long double logl.
// Usage: math_check N SEED -> one line of key=value pairs per input range: the largest
// error in units in the last place of the double result and the count of results that
// are not the correctly rounded ones; then the special values and the powers of two.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>

#include "gf_math.h"

static double ulp_of(double v) {
    v = std::fabs(v);
    if(v < 0x1p-1022)
        return 0x1p-1074;
    return std::nextafter(v, HUGE_VAL) - v;
}

struct Stat {
    double max_ulp = 0.0, worst_x = 0.0;
    long long n = 0, not_nearest = 0;
    void add(double x, double got, long double ref) {
        const double nearest = (double)ref;
        n++;
        if(got != nearest)
            not_nearest++;
        if(nearest == 0.0 || std::isinf(nearest))
            return;
        const double e = (double)(std::fabs((long double)got - ref) / (long double)ulp_of(nearest));
        if(e > max_ulp) {
            max_ulp = e;
            worst_x = x;
        }
    }
    void print(const char *name) const {
        std::printf("%s n=%lld max_ulp=%.4f worst_x=%.17g not_nearest=%lld\n", name, n, max_ulp, worst_x,
                    not_nearest);
    }
};

int main(int argc, char **argv) {
    const long long N = argc > 1 ? std::atoll(argv[1]) : 1000000;
    const unsigned long long seed = argc > 2 ? std::strtoull(argv[2], nullptr, 10) : 1;
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> u(0.0, 1.0);

    Stat all, near1, cancel, sub;
    for(long long i = 0; i < N; i++) {
        double x = std::ldexp(0.5 + 0.5 * u(rng), -1021 + (int)(u(rng) * 2044.0));  // normal x, every exponent
        all.add(x, gfmath::log_d(x), logl((long double)x));
        x = 0.5 + 1.5 * u(rng);  // around 1, where log x is small
        near1.add(x, gfmath::log_d(x), logl((long double)x));
        // x in [sqrt(2)/4, sqrt(2)/2) and [sqrt(2), 2 sqrt(2)): e ln 2 and log m cancel
        x = (u(rng) < 0.5 ? 0.35355339059327373 : 1.4142135623730951) * (1.0 + u(rng));
        cancel.add(x, gfmath::log_d(x), logl((long double)x));
        x = std::ldexp(u(rng), -1022 - (int)(u(rng) * 52.0));  // subnormal x
        if(x > 0.0)
            sub.add(x, gfmath::log_d(x), logl((long double)x));
    }
    all.print("log_range");
    near1.print("log_near_one");
    cancel.print("log_cancelling");
    sub.print("log_subnormal_inputs");

    const double xs[] = {0.0, -0.0, 1.0, -1.0, HUGE_VAL, -HUGE_VAL, NAN, 0x1p-1074, 2.0, 0.5, 0x1.fffffffffffffp1023};
    for(double x : xs)
        std::printf("special x=%.17g log=%.17g\n", x, gfmath::log_d(x));
    long long pow2_bad = 0;
    for(int k = -1074; k <= 1023; k++) {
        const double x = std::ldexp(1.0, k);
        if(gfmath::log_d(x) != (double)((long double)k * logl(2.0L)))
            pow2_bad++;
    }
    std::printf("log_powers_of_two not_nearest=%lld\n", pow2_bad);

    // the sum of -log(v) as one logarithm of a frexp-normalised product, against the
    // long double sum of -logl(v) (Kahan-compensated), over values spread across every
    // binary exponent with subnormals, zeros, negatives and NaNs mixed in (which add 0)
    for(int trial = 0; trial < 3; trial++) {
        const long long n = trial == 0 ? N : (trial == 1 ? 4 : 1000);
        double m = 1.0, e = 0.0;
        long double ref = 0.0L, comp = 0.0L, scale = 0.0L;
        for(long long i = 0; i < n; i++) {
            const double r = u(rng);
            double v;
            if(trial == 1)
                v = std::ldexp(0.5 + 0.5 * u(rng), -10 + (int)(u(rng) * 17.0));  // likelihood-like values
            else if(r < 0.05)
                v = 0.0;
            else if(r < 0.07)
                v = -u(rng);
            else if(r < 0.08)
                v = NAN;
            else
                v = std::ldexp(u(rng), 1023 - (int)(u(rng) * 2098.0));  // 2^1023 down to subnormals
            gfmath::nll_add(m, e, v);
            if(v > 0.0) {
                const long double t = -logl((long double)v) - comp;
                const long double s = ref + t;
                comp                = (s - ref) - t;
                ref                 = s;
                scale += fabsl(logl((long double)v));
            }
        }
        const double got = gfmath::nll_total(m, e);
        std::printf("nll_product trial=%d n=%lld rel_err=%.3e ulps_of_scale=%.3f\n", trial, n,
                    (double)(fabsl((long double)got - ref) / fabsl(ref)),
                    (double)(fabsl((long double)got - ref) / (scale * 0x1p-53L)));
    }
    {
        double m = 1.0, e = 0.0;
        const double got_empty = gfmath::nll_total(m, e);
        gfmath::nll_add(m, e, 0.25);
        gfmath::nll_add(m, e, HUGE_VAL);
        gfmath::nll_add(m, e, 3.0);
        std::printf("nll_special empty=%.17g with_inf=%.17g\n", got_empty, gfmath::nll_total(m, e));
        double m2 = 1.0, e2 = 0.0;
        gfmath::nll_add(m2, e2, 0.125);
        gfmath::nll_add(m2, e2, 8.0);
        std::printf("nll_special exact_pair=%.17g\n", gfmath::nll_total(m2, e2));
    }
    {
        // the integer-exponent form against the double form, factor by factor
        std::mt19937_64 rng(77);
        std::uniform_real_distribution<double> u(0.0, 1.0);
        double m = 1.0, e = 0.0, mi = 1.0;
        int ei = 0;
        long long bad = 0, n = 0;
        for(; n < 1000000; n++) {
            const double r = u(rng);
            const double v = r < 0.02 ? 0.0 : r < 0.03 ? NAN : std::ldexp(u(rng), 1023 - (int)(u(rng) * 2098.0));
            gfmath::nll_add(m, e, v);
            gfmath::nll_add(mi, ei, v);
            if(std::memcmp(&m, &mi, sizeof m) != 0 || e != (double)ei)
                bad++;
        }
        std::printf("nll_int_exponent n=%lld mismatches=%lld\n", n, bad);
    }
    return 0;
}
