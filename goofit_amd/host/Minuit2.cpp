// Minuit2.cpp -- variable-metric minimiser behind the Minuit2 API (include/Minuit2/Minuit2.h).
//
// Method (F. James, "MINUIT: Function Minimization and Error Analysis"):
//   * parameters with limits are minimised in internal coordinates
//     (sin transform for two limits, square-root transforms for one);
//   * gradients by central differences with per-parameter steps chosen from
//     the second-derivative estimate and the function's rounding level;
//   * a Davidon-Fletcher-Powell update of the inverse Hessian V, switched to
//     the BFGS form when dx.dg > dg.V.dg (Fletcher's switch);
//   * a parabolic line search along -V g;
//   * convergence when EDM = g.V.g / 2 < 0.002 * tolerance * Up; with strategy
//     >= 1 the inverse Hessian is then recomputed from numerical second
//     derivatives (HESSE) and the iteration resumes if the EDM is still large.
// Errors: covariance = 2 Up J V J^T with J the Jacobian of the internal-to-external
// transform at the minimum.
#include <Minuit2/Minuit2.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <iomanip>
#include <iostream>
#include <limits>
#include <stdexcept>

namespace ROOT {
namespace Minuit2 {

namespace {

int g_print_level = 0;

constexpr double kEps  = std::numeric_limits<double>::epsilon();
const double kEps2     = 2.0 * std::sqrt(kEps);

// ------------------------------------------------------------ transforms

struct Trans {
    int ext    = 0;  // external index
    int kind   = 0;  // 0 none, 1 two limits (sin), 2 lower (sqrt), 3 upper (sqrt)
    double lo  = 0, up = 0;

    double int2ext(double x) const {
        switch(kind) {
        case 1: return lo + 0.5 * (up - lo) * (std::sin(x) + 1.0);
        case 2: return lo - 1.0 + std::sqrt(x * x + 1.0);
        case 3: return up + 1.0 - std::sqrt(x * x + 1.0);
        default: return x;
        }
    }
    double ext2int(double v) const {
        switch(kind) {
        case 1: {
            double a = 2.0 * (v - lo) / (up - lo) - 1.0;
            a        = std::max(-1.0, std::min(1.0, a));
            return std::asin(a);
        }
        case 2: {
            double t = v - lo + 1.0;
            return t > 1.0 ? std::sqrt(t * t - 1.0) : 0.0;
        }
        case 3: {
            double t = up - v + 1.0;
            return t > 1.0 ? std::sqrt(t * t - 1.0) : 0.0;
        }
        default: return v;
        }
    }
    double dext(double x) const {
        switch(kind) {
        case 1: return 0.5 * (up - lo) * std::cos(x);
        case 2: return x / std::sqrt(x * x + 1.0);
        case 3: return -x / std::sqrt(x * x + 1.0);
        default: return 1.0;
        }
    }
    // internal step equivalent to an external error
    double int_error(double v, double err) const {
        if(kind == 0)
            return err;
        double x  = ext2int(v);
        double x1 = ext2int(v + err), x2 = ext2int(v - err);
        double d  = 0.5 * (std::fabs(x1 - x) + std::fabs(x2 - x));
        if(!(d > 0))
            d = std::fabs(x1 - x) + std::fabs(x2 - x);
        if(kind == 1)
            d = std::min(d, 1.0);
        return d > 0 ? d : 0.1;
    }
};

// The function in internal coordinates of the free parameters.
struct IntFcn {
    const FCNBase &fcn;
    std::vector<MinuitParameter> pars;
    std::vector<Trans> tr;
    mutable unsigned int nfcn = 0;

    IntFcn(const FCNBase &f, const MnUserParameters &p)
        : fcn(f)
        , pars(p.Parameters()) {
        for(size_t i = 0; i < pars.size(); i++) {
            if(pars[i].fixed)
                continue;
            Trans t;
            t.ext = (int)i;
            if(pars[i].has_lower && pars[i].has_upper) {
                t.kind = 1;
                t.lo   = pars[i].lower;
                t.up   = pars[i].upper;
            } else if(pars[i].has_lower) {
                t.kind = 2;
                t.lo   = pars[i].lower;
            } else if(pars[i].has_upper) {
                t.kind = 3;
                t.up   = pars[i].upper;
            }
            tr.push_back(t);
        }
    }
    size_t n() const { return tr.size(); }
    std::vector<double> ext(const std::vector<double> &x) const {
        std::vector<double> v(pars.size());
        for(size_t i = 0; i < pars.size(); i++)
            v[i] = pars[i].value;
        for(size_t k = 0; k < tr.size(); k++)
            v[tr[k].ext] = tr[k].int2ext(x[k]);
        return v;
    }
    double operator()(const std::vector<double> &x) const {
        nfcn++;
        return fcn(ext(x));
    }
    // independent points in one request (FCNBase::EvaluateMany)
    std::vector<double> many(const std::vector<std::vector<double>> &xs) const {
        if(xs.empty())
            return {};
        nfcn += (unsigned int)xs.size();
        std::vector<std::vector<double>> e;
        e.reserve(xs.size());
        for(const auto &x : xs)
            e.push_back(ext(x));
        std::vector<double> f = fcn.EvaluateMany(e);
        if(f.size() != xs.size())
            throw std::runtime_error("FCNBase::EvaluateMany returned a wrong number of values");
        return f;
    }
    std::vector<double> start() const {
        std::vector<double> x(tr.size());
        for(size_t k = 0; k < tr.size(); k++)
            x[k] = tr[k].ext2int(pars[tr[k].ext].value);
        return x;
    }
    std::vector<double> start_errors() const {
        std::vector<double> e(tr.size());
        for(size_t k = 0; k < tr.size(); k++) {
            double err = pars[tr[k].ext].error;
            if(!(err > 0))
                err = 0.1 * std::max(std::fabs(pars[tr[k].ext].value), 1.0);
            e[k] = tr[k].int_error(pars[tr[k].ext].value, err);
        }
        return e;
    }
};

// ------------------------------------------------------------ linear algebra

using Mat = std::vector<double>;  // row-major n x n

// Cholesky inverse of a symmetric matrix; false if not positive definite.
bool spd_inverse(const Mat &A, size_t n, Mat &inv) {
    Mat L(n * n, 0.0);
    for(size_t i = 0; i < n; i++)
        for(size_t j = 0; j <= i; j++) {
            double s = A[i * n + j];
            for(size_t k = 0; k < j; k++)
                s -= L[i * n + k] * L[j * n + k];
            if(i == j) {
                if(!(s > 0))
                    return false;
                L[i * n + i] = std::sqrt(s);
            } else {
                L[i * n + j] = s / L[j * n + j];
            }
        }
    // inverse of L
    Mat Li(n * n, 0.0);
    for(size_t i = 0; i < n; i++) {
        Li[i * n + i] = 1.0 / L[i * n + i];
        for(size_t j = 0; j < i; j++) {
            double s = 0;
            for(size_t k = j; k < i; k++)
                s -= L[i * n + k] * Li[k * n + j];
            Li[i * n + j] = s / L[i * n + i];
        }
    }
    inv.assign(n * n, 0.0);
    for(size_t i = 0; i < n; i++)
        for(size_t j = 0; j <= i; j++) {
            double s = 0;
            for(size_t k = i; k < n; k++)
                s += Li[k * n + i] * Li[k * n + j];
            inv[i * n + j] = inv[j * n + i] = s;
        }
    return true;
}

// Make a symmetric matrix positive definite by shifting its diagonal (MnPosDef idea).
bool make_pos_def(Mat &A, size_t n) {
    Mat tmp;
    if(spd_inverse(A, n, tmp))
        return false;
    double dmin = 1e300, dmax = 0;
    for(size_t i = 0; i < n; i++) {
        dmin = std::min(dmin, A[i * n + i]);
        dmax = std::max(dmax, std::fabs(A[i * n + i]));
    }
    double shift = (dmin <= 0 ? -dmin : 0) + 1e-3 * std::max(dmax, 1e-10);
    for(int it = 0; it < 60; it++) {
        Mat B = A;
        for(size_t i = 0; i < n; i++)
            B[i * n + i] += shift;
        if(spd_inverse(B, n, tmp)) {
            A = B;
            return true;
        }
        shift *= 2;
    }
    // fall back to the diagonal
    for(size_t i = 0; i < n; i++)
        for(size_t j = 0; j < n; j++)
            if(i != j)
                A[i * n + j] = 0;
    for(size_t i = 0; i < n; i++)
        A[i * n + i] = std::max(std::fabs(A[i * n + i]), 1e-10);
    return true;
}

double dot(const std::vector<double> &a, const std::vector<double> &b) {
    double s = 0;
    for(size_t i = 0; i < a.size(); i++)
        s += a[i] * b[i];
    return s;
}

std::vector<double> matvec(const Mat &A, const std::vector<double> &x) {
    size_t n = x.size();
    std::vector<double> y(n, 0.0);
    for(size_t i = 0; i < n; i++) {
        double s = 0;
        for(size_t j = 0; j < n; j++)
            s += A[i * n + j] * x[j];
        y[i] = s;
    }
    return y;
}

// ------------------------------------------------------------ derivatives

struct Grad {
    std::vector<double> g, g2, step;
};

// Central-difference gradient with step refinement (cf. Minuit2's numerical
// two-point gradient): the step of parameter i aims at sqrt(dfmin / |g2_i|),
// dfmin being the rounding level 8 eps2 (|f| + Up) of the function value.
// Each parameter's refinement depends only on its own two function values, so
// the cycles run parameter-parallel: cycle c evaluates f(x +- step_i e_i) of
// every parameter still refining as one request of 2 n_active points.  Given a
// function without side effects this is the parameter-by-parameter loop.
Grad numerical_gradient(const IntFcn &F, const std::vector<double> &x, double f0, const Grad &prev, double up,
                        unsigned int ncycles) {
    const size_t n = x.size();
    Grad G;
    G.g.assign(n, 0);
    G.g2.assign(n, 0);
    G.step.assign(n, 0);
    const double dfmin  = 8.0 * kEps2 * (std::fabs(f0) + up);
    const double vrysml = 8.0 * kEps * kEps;
    std::vector<double> step(n), stepb4(n, 0.0), g(n), g2(n);
    std::vector<char> active(n, 1);
    for(size_t i = 0; i < n; i++) {
        step[i] = prev.step[i];
        g[i]    = prev.g[i];
        g2[i]   = prev.g2[i];
    }
    std::vector<size_t> idx;
    std::vector<std::vector<double>> pts;
    for(unsigned int c = 0; c < ncycles; c++) {
        idx.clear();
        pts.clear();
        for(size_t i = 0; i < n; i++) {
            if(!active[i])
                continue;
            const double xtf    = x[i];
            const double epspri = kEps2 + std::fabs(prev.g[i] * kEps2);
            double optstp       = std::sqrt(dfmin / (std::fabs(g2[i]) + epspri));
            double st           = std::max(optstp, std::fabs(0.1 * step[i]));
            double stpmax       = 10.0 * std::fabs(prev.step[i] > 0 ? prev.step[i] : st);
            if(F.tr[i].kind == 1)
                stpmax = std::min(stpmax, 0.5);
            double stpmin = std::max(vrysml, 8.0 * std::fabs(kEps2 * xtf));
            st            = std::min(st, stpmax);
            st            = std::max(st, stpmin);
            step[i]       = st;
            if(c > 0 && std::fabs((st - stepb4[i]) / st) < 0.3) {
                active[i] = 0;
                continue;
            }
            stepb4[i] = st;
            idx.push_back(i);
            std::vector<double> xp = x, xm = x;
            xp[i] = xtf + st;
            xm[i] = xtf - st;
            pts.push_back(std::move(xp));
            pts.push_back(std::move(xm));
        }
        if(idx.empty())
            break;
        const std::vector<double> f = F.many(pts);
        for(size_t k = 0; k < idx.size(); k++) {
            const size_t i     = idx[k];
            const double fs1   = f[2 * k], fs2 = f[2 * k + 1];
            const double grdb4 = g[i];
            g[i]               = 0.5 * (fs1 - fs2) / step[i];
            g2[i]              = (fs1 + fs2 - 2.0 * f0) / step[i] / step[i];
            if(c > 0 && std::fabs(grdb4 - g[i]) / (std::fabs(g[i]) + dfmin / step[i]) < 0.05)
                active[i] = 0;
        }
    }
    for(size_t i = 0; i < n; i++) {
        G.g[i]    = g[i];
        G.g2[i]   = g2[i];
        G.step[i] = step[i];
    }
    return G;
}

// Full numerical Hessian (cf. MnHesse): diagonal from f(x +- d), off-diagonal
// from the symmetric pair f(x + d_i + d_j), f(x - d_i - d_j) together with the
// diagonal's f(x +- d_i), f(x +- d_j):
//     H_ij = [f(++) + f(--) - f(+i) - f(-i) - f(+j) - f(-j) + 2 f0] / (2 d_i d_j),
// whose error is O(d^2).  MnHesse's one-sided f(x + d_i + d_j) formula has an
// O(d) error from the third derivatives; with steps of about 0.3 sigma and
// correlations near 0.97 (Dalitz couplings) that moved single errors by up to
// 15%.  Steps aim at a function change of about 0.1 Up.
bool numerical_hessian_once(const IntFcn &F, const std::vector<double> &x, double f0, double up, const Grad &grad,
                            double target, Mat &H, std::vector<double> &g_out, std::vector<double> &g2_out,
                            std::vector<double> &d_out) {
    const size_t n = x.size();
    H.assign(n * n, 0.0);
    g_out.assign(n, 0.0);
    g2_out.assign(n, 0.0);
    d_out.assign(n, 0.0);
    std::vector<double> fp(n, 0.0), fm(n, 0.0);
    const double dfmin = 8.0 * kEps2 * (std::fabs(f0) + up);
    // diagonal: each parameter's step search (up to 6 tries) depends only on its
    // own two values, so the tries run parameter-parallel, one request per round
    std::vector<double> d(n), fs1(n, 0.0), fs2(n, 0.0);
    std::vector<char> active(n, 1);
    for(size_t i = 0; i < n; i++) {
        double g2 = grad.g2[i];
        d[i] = (std::fabs(g2) > 0) ? std::sqrt(2.0 * target * up / std::fabs(g2)) : std::fabs(grad.step[i]) * 10;
        if(!(d[i] > 0))
            d[i] = 1e-3;
        if(F.tr[i].kind == 1)
            d[i] = std::min(d[i], 0.5);
    }
    std::vector<size_t> idx;
    std::vector<std::vector<double>> pts;
    for(int c = 0; c < 6; c++) {
        idx.clear();
        pts.clear();
        for(size_t i = 0; i < n; i++) {
            if(!active[i])
                continue;
            idx.push_back(i);
            std::vector<double> xp = x, xm = x;
            xp[i] = x[i] + d[i];
            xm[i] = x[i] - d[i];
            pts.push_back(std::move(xp));
            pts.push_back(std::move(xm));
        }
        if(idx.empty())
            break;
        const std::vector<double> f = F.many(pts);
        for(size_t k = 0; k < idx.size(); k++) {
            const size_t i = idx[k];
            fs1[i]         = f[2 * k];
            fs2[i]         = f[2 * k + 1];
            double sag     = 0.5 * (fs1[i] + fs2[i] - 2.0 * f0);
            if(sag > 4.0 * dfmin && sag < 100.0 * target * up) {
                active[i] = 0;
                continue;
            }
            if(c == 5) {
                active[i] = 0;  // the last try keeps its values and step
                continue;
            }
            if(sag <= 4.0 * dfmin)
                d[i] *= 4.0;
            else
                d[i] *= 0.3;
        }
    }
    for(size_t i = 0; i < n; i++) {
        H[i * n + i] = (fs1[i] + fs2[i] - 2.0 * f0) / (d[i] * d[i]);
        g_out[i]     = 0.5 * (fs1[i] - fs2[i]) / d[i];
        g2_out[i]    = H[i * n + i];
        d_out[i]     = d[i];
        fp[i]        = fs1[i];
        fm[i]        = fs2[i];
    }
    // off-diagonal: every f(x +- (d_i e_i + d_j e_j)) in one request
    pts.clear();
    for(size_t i = 0; i < n; i++)
        for(size_t j = i + 1; j < n; j++) {
            std::vector<double> xp = x, xm = x;
            xp[i] = x[i] + d_out[i];
            xp[j] = x[j] + d_out[j];
            xm[i] = x[i] - d_out[i];
            xm[j] = x[j] - d_out[j];
            pts.push_back(std::move(xp));
            pts.push_back(std::move(xm));
        }
    const std::vector<double> fo = F.many(pts);
    size_t k = 0;
    for(size_t i = 0; i < n; i++)
        for(size_t j = i + 1; j < n; j++, k += 2) {
            double h = (fo[k] + fo[k + 1] - fp[i] - fm[i] - fp[j] - fm[j] + 2.0 * f0) / (2.0 * d_out[i] * d_out[j]);
            H[i * n + j] = H[j * n + i] = h;
        }
    return true;
}

// Hessian with steps aimed at a function change of 0.1 Up, refined with smaller
// steps while the estimate is not positive definite (curved valleys).
bool numerical_hessian(const IntFcn &F, const std::vector<double> &x, double f0, double up, const Grad &grad,
                       Mat &H, std::vector<double> &g_out, std::vector<double> &g2_out, std::vector<double> &d_out) {
    double target = 0.1;
    for(int a = 0; a < 4; a++) {
        numerical_hessian_once(F, x, f0, up, grad, target, H, g_out, g2_out, d_out);
        Mat inv;
        if(spd_inverse(H, x.size(), inv))
            return true;
        target *= 0.01;
    }
    return false;
}

// ------------------------------------------------------------ line search

// Parabolic search of min_a f(x + a*dir), a in (0, 4], starting at a = 1.
double line_search(const IntFcn &F, const std::vector<double> &x, const std::vector<double> &dir, double f0,
                   double slope, double &fbest, double up) {
    auto at = [&](double a) {
        std::vector<double> y(x.size());
        for(size_t i = 0; i < x.size(); i++)
            y[i] = x[i] + a * dir[i];
        return F(y);
    };
    const double toler = 0.05, amax = 4.0;
    double a1 = 1.0, f1 = at(a1);
    double abest = 0.0;
    fbest        = f0;
    if(f1 < fbest) {
        fbest = f1;
        abest = a1;
    }
    // parabola through (0, f0, slope) and (a1, f1)
    double a2 = 0, f2 = 0;
    for(int it = 0; it < 12; it++) {
        double denom = f1 - f0 - slope * a1;
        double anew;
        if(denom <= 0)
            anew = std::min(amax, 2.0 * a1);  // no curvature: expand
        else
            anew = -slope * a1 * a1 / (2.0 * denom);
        anew = std::max(anew, 0.02 * a1);
        anew = std::min(anew, amax);
        if(std::fabs(anew - a1) < toler * std::max(a1, 1e-3) && f1 <= f0 + slope * a1 * 1e-4)
            break;
        a2 = anew;
        f2 = at(a2);
        if(f2 < fbest) {
            fbest = f2;
            abest = a2;
        }
        // converged when the new point changes the minimum little
        if(std::fabs(f2 - f1) < 1e-3 * up && fbest < f0)
            break;
        // keep the better of the two as the next interpolation point
        if(f2 < f1 || denom <= 0) {
            a1 = a2;
            f1 = f2;
        } else {
            // refine with the three-point parabola (0, a2, a1)
            double x0 = 0, x1 = a2, x2 = a1, y0 = f0, y1 = f2, y2 = f1;
            double d   = (x1 - x0) * (x2 - x0) * (x2 - x1);
            if(d == 0)
                break;
            double A = ((y2 - y0) * (x1 - x0) - (y1 - y0) * (x2 - x0)) / d;
            double B = ((y1 - y0) * (x2 - x0) * (x2 - x0) - (y2 - y0) * (x1 - x0) * (x1 - x0)) / d;
            if(A <= 0) {
                a1 = a2;
                f1 = f2;
                continue;
            }
            double a3 = -B / (2 * A);
            a3        = std::max(0.02 * a2, std::min(a3, amax));
            double f3 = at(a3);
            if(f3 < fbest) {
                fbest = f3;
                abest = a3;
            }
            if(std::fabs(a3 - a2) < toler * a2)
                break;
            a1 = a3;
            f1 = f3;
        }
        if(abest > 0 && it >= 2)
            break;
    }
    return abest;
}

struct MigradResult {
    std::vector<double> x, g, g2, step;
    Mat V;
    double f = 0, edm = 0;
    bool converged = false, call_limit = false, made_pos_def = false, hesse_ok = false, hesse_failed = false;
};

void print_iter(int it, double f, double edm, unsigned int nfcn) {
    if(g_print_level >= 2)
        std::cout << "MIGRAD iteration " << it << "  FCN = " << std::setprecision(12) << f << "  Edm = " << edm
                  << "  NCalls = " << nfcn << std::endl;
}

// Run the variable-metric iterations from x (internal coordinates).
MigradResult migrad_core(const IntFcn &F, std::vector<double> x, const std::vector<double> &err0, unsigned int maxfcn,
                         double edmval, const MnStrategy &strat, double up) {
    const size_t n = x.size();
    MigradResult R;
    R.x = x;
    if(n == 0) {
        R.f         = F(x);
        R.converged = true;
        R.hesse_ok  = true;
        return R;
    }
    double f = F(x);
    Grad prev;
    prev.g.assign(n, 0.0);
    prev.g2.assign(n, 0.0);
    prev.step.assign(n, 0.0);
    for(size_t i = 0; i < n; i++) {
        double e      = err0[i];
        prev.step[i]  = 0.1 * e;
        prev.g2[i]    = 2.0 * up / (e * e);  // a function rising by Up at one error
        prev.g[i]     = 0;
    }
    Grad G = numerical_gradient(F, x, f, prev, up, strat.GradientNCycles());
    Mat V(n * n, 0.0);
    for(size_t i = 0; i < n; i++)
        V[i * n + i] = (G.g2[i] > 0) ? 1.0 / G.g2[i] : (err0[i] * err0[i] / (2.0 * up));
    double edm       = 0.5 * dot(G.g, matvec(V, G.g));
    bool hesse_done  = false;
    int iter         = 0;
    for(;;) {
        print_iter(iter, f, edm, F.nfcn);
        if(F.nfcn >= maxfcn) {
            R.call_limit = true;
            break;
        }
        if(edm < edmval && edm >= 0) {
            if(strat.Strategy() >= 1 && !hesse_done) {
                Mat H;
                std::vector<double> hg, hg2, hd;
                numerical_hessian(F, x, f, up, G, H, hg, hg2, hd);
                Mat Vh;
                hesse_done = true;
                if(spd_inverse(H, n, Vh)) {
                    V              = Vh;
                    R.hesse_ok     = true;
                    R.hesse_failed = false;
                    R.made_pos_def = false;  // the final error matrix is the true inverse Hessian
                } else {
                    R.hesse_ok = false;
                    R.made_pos_def = make_pos_def(H, n);
                    if(spd_inverse(H, n, Vh))
                        V = Vh;
                    R.hesse_failed = true;
                }
                G.g2 = hg2;
                for(size_t i = 0; i < n; i++)
                    if(hd[i] > 0)
                        G.step[i] = std::min(G.step[i] > 0 ? G.step[i] : hd[i], hd[i]);
                edm = 0.5 * dot(G.g, matvec(V, G.g));
                if(edm < edmval) {
                    R.converged = true;
                    break;
                }
                continue;  // resume the iterations with the accurate V
            }
            R.converged = true;
            break;
        }
        // search direction
        std::vector<double> step = matvec(V, G.g);
        for(double &s : step)
            s = -s;
        double gdel = dot(step, G.g);
        if(gdel >= 0) {
            // not a descent direction: make V positive definite and retry
            R.made_pos_def = make_pos_def(V, n) || R.made_pos_def;
            step           = matvec(V, G.g);
            for(double &s : step)
                s = -s;
            gdel = dot(step, G.g);
            if(gdel >= 0) {
                for(size_t i = 0; i < n; i++)
                    step[i] = -G.g[i] * (G.g2[i] > 0 ? 1.0 / G.g2[i] : 1.0);
                gdel = dot(step, G.g);
            }
        }
        double fnew  = f;
        double alpha = line_search(F, x, step, f, gdel, fnew, up);
        if(alpha <= 0 || !(fnew < f)) {
            // no progress along the direction: refresh derivatives once, else stop
            if(!hesse_done) {
                Mat H;
                std::vector<double> hg, hg2, hd;
                numerical_hessian(F, x, f, up, G, H, hg, hg2, hd);
                hesse_done = true;
                Mat Vh;
                if(!spd_inverse(H, n, Vh)) {
                    R.made_pos_def = make_pos_def(H, n);
                    spd_inverse(H, n, Vh);
                }
                V    = Vh;
                G.g  = hg;
                G.g2 = hg2;
                edm  = 0.5 * dot(G.g, matvec(V, G.g));
                iter++;
                continue;
            }
            break;
        }
        std::vector<double> xnew(n);
        for(size_t i = 0; i < n; i++)
            xnew[i] = x[i] + alpha * step[i];
        Grad Gn = numerical_gradient(F, xnew, fnew, G, up, strat.GradientNCycles());
        // DFP update with Fletcher's switch to BFGS
        std::vector<double> dx(n), dg(n);
        for(size_t i = 0; i < n; i++) {
            dx[i] = xnew[i] - x[i];
            dg[i] = Gn.g[i] - G.g[i];
        }
        double delgam          = dot(dx, dg);
        std::vector<double> vg = matvec(V, dg);
        double gvg             = dot(dg, vg);
        if(delgam > 0 && gvg > 0) {
            for(size_t i = 0; i < n; i++)
                for(size_t j = 0; j < n; j++)
                    V[i * n + j] += dx[i] * dx[j] / delgam - vg[i] * vg[j] / gvg;
            if(delgam > gvg) {
                std::vector<double> flnu(n);
                for(size_t i = 0; i < n; i++)
                    flnu[i] = dx[i] / delgam - vg[i] / gvg;
                for(size_t i = 0; i < n; i++)
                    for(size_t j = 0; j < n; j++)
                        V[i * n + j] += gvg * flnu[i] * flnu[j];
            }
        }
        x   = xnew;
        f   = fnew;
        G   = Gn;
        edm = 0.5 * dot(G.g, matvec(V, G.g));
        if(edm < 0) {
            R.made_pos_def = make_pos_def(V, n) || R.made_pos_def;
            edm            = 0.5 * dot(G.g, matvec(V, G.g));
        }
        hesse_done = false;
        iter++;
    }
    R.x    = x;
    R.f    = f;
    R.g    = G.g;
    R.g2   = G.g2;
    R.step = G.step;
    R.V    = V;
    R.edm  = edm;
    return R;
}

void fill_state(const IntFcn &F, const MigradResult &R, double up, MnUserParameterState &st) {
    const size_t n  = F.n();
    const size_t ne = F.pars.size();
    st.pars_        = MnUserParameters();
    st.pars_        = MnUserParameters();
    // rebuild the user parameters with the new values and errors
    MnUserParameters out;
    for(size_t i = 0; i < ne; i++) {
        const MinuitParameter &p = F.pars[i];
        if(p.fixed)
            out.Add(p.name, p.value);
        else if(p.has_lower && p.has_upper)
            out.Add(p.name, p.value, p.error, p.lower, p.upper);
        else
            out.Add(p.name, p.value, p.error);
        if(!p.fixed && (p.has_lower != p.has_upper)) {
            if(p.has_lower)
                out.SetLowerLimit((unsigned int)i, p.lower);
            else
                out.SetUpperLimit((unsigned int)i, p.upper);
        }
    }
    std::vector<double> J(n);
    for(size_t k = 0; k < n; k++) {
        out.SetValue(F.tr[k].ext, F.tr[k].int2ext(R.x[k]));
        J[k] = F.tr[k].dext(R.x[k]);
    }
    st.cov_.assign(ne * ne, 0.0);
    if(!R.V.empty()) {
        for(size_t a = 0; a < n; a++)
            for(size_t b = 0; b < n; b++)
                st.cov_[F.tr[a].ext * ne + F.tr[b].ext] = 2.0 * up * J[a] * R.V[a * n + b] * J[b];
        for(size_t k = 0; k < n; k++)
            out.SetError(F.tr[k].ext, std::sqrt(std::max(st.cov_[F.tr[k].ext * ne + F.tr[k].ext], 0.0)));
    }
    st.pars_ = out;
    st.fval_ = R.f;
    st.edm_  = R.edm;
    st.nfcn_ = F.nfcn;
}

}  // namespace

// ------------------------------------------------------------ MnUserParameters

bool MnUserParameters::Add(const std::string &name, double val) {
    for(const auto &p : pars_)
        if(p.name == name)
            return false;
    MinuitParameter p;
    p.name  = name;
    p.value = val;
    p.fixed = true;
    pars_.push_back(p);
    return true;
}

bool MnUserParameters::Add(const std::string &name, double val, double err) {
    for(const auto &p : pars_)
        if(p.name == name)
            return false;
    MinuitParameter p;
    p.name  = name;
    p.value = val;
    p.error = err;
    pars_.push_back(p);
    return true;
}

bool MnUserParameters::Add(const std::string &name, double val, double err, double low, double up) {
    for(const auto &p : pars_)
        if(p.name == name)
            return false;
    MinuitParameter p;
    p.name  = name;
    p.value = val;
    p.error = err;
    if(low == up) {
        pars_.push_back(p);
        return true;
    }
    p.lower     = std::min(low, up);
    p.upper     = std::max(low, up);
    p.has_lower = p.has_upper = true;
    pars_.push_back(p);
    return true;
}

unsigned int MnUserParameters::Index(const std::string &name) const {
    for(size_t i = 0; i < pars_.size(); i++)
        if(pars_[i].name == name)
            return (unsigned int)i;
    throw std::out_of_range("Minuit2: no parameter named " + name);
}

void MnUserParameters::SetLimits(unsigned int i, double low, double up) {
    MinuitParameter &p = pars_.at(i);
    p.lower            = std::min(low, up);
    p.upper            = std::max(low, up);
    p.has_lower = p.has_upper = true;
}
void MnUserParameters::SetLowerLimit(unsigned int i, double low) {
    pars_.at(i).lower     = low;
    pars_.at(i).has_lower = true;
    pars_.at(i).has_upper = false;
}
void MnUserParameters::SetUpperLimit(unsigned int i, double up) {
    pars_.at(i).upper     = up;
    pars_.at(i).has_upper = true;
    pars_.at(i).has_lower = false;
}
void MnUserParameters::RemoveLimits(unsigned int i) { pars_.at(i).has_lower = pars_.at(i).has_upper = false; }

unsigned int MnUserParameters::VariableParameters() const {
    unsigned int n = 0;
    for(const auto &p : pars_)
        n += p.fixed ? 0 : 1;
    return n;
}
std::vector<double> MnUserParameters::Params() const {
    std::vector<double> v;
    for(const auto &p : pars_)
        v.push_back(p.value);
    return v;
}
std::vector<double> MnUserParameters::Errors() const {
    std::vector<double> v;
    for(const auto &p : pars_)
        v.push_back(p.error);
    return v;
}

double MnUserParameterState::Covariance(unsigned int i, unsigned int j) const {
    size_t n = pars_.NParameters();
    if(cov_.size() != n * n)
        return 0.0;
    return cov_.at(i * n + j);
}

// ------------------------------------------------------------ MnMigrad

FunctionMinimum MnMigrad::operator()(unsigned int maxfcn, double tolerance) {
    IntFcn F(fcn_, par_);
    const size_t n = F.n();
    if(maxfcn == 0)
        maxfcn = (unsigned int)(200 + 100 * n + 5 * n * n);
    const double up     = fcn_.Up();
    const double edmval = 0.002 * tolerance * up;
    MigradResult R      = migrad_core(F, F.start(), F.start_errors(), maxfcn, edmval, strategy_, up);
    FunctionMinimum M;
    M.up_ = up;
    fill_state(F, R, up, M.state_);
    M.has_cov_       = !R.V.empty() || n == 0;
    M.pos_def_       = !R.made_pos_def;
    M.made_pos_def_  = R.made_pos_def;
    M.hesse_failed_  = R.hesse_failed;
    M.accurate_      = R.hesse_ok && !R.made_pos_def;
    M.call_limit_    = R.call_limit;
    M.above_max_edm_ = !(R.edm < edmval);
    M.valid_params_  = std::isfinite(R.f);
    M.valid_         = R.converged && !R.call_limit && std::isfinite(R.f) && !R.made_pos_def && !M.above_max_edm_;
    M.state_.valid_  = M.valid_;
    M.x_int_         = R.x;
    M.v_int_         = R.V;
    if(g_print_level >= 1)
        std::cout << M;
    return M;
}

// ------------------------------------------------------------ MnHesse

MnUserParameterState MnHesse::operator()(const FCNBase &fcn, const MnUserParameters &par, unsigned int) const {
    IntFcn F(fcn, par);
    const size_t n   = F.n();
    const double up  = fcn.Up();
    std::vector<double> x = F.start(), e = F.start_errors();
    double f0 = F(x);
    Grad G;
    G.g.assign(n, 0);
    G.g2.assign(n, 0);
    G.step.assign(n, 0);
    for(size_t i = 0; i < n; i++) {
        G.g2[i]   = 2.0 * up / (e[i] * e[i]);
        G.step[i] = 0.1 * e[i];
    }
    Mat H;
    std::vector<double> hg, hg2, hd;
    numerical_hessian(F, x, f0, up, G, H, hg, hg2, hd);
    MigradResult R;
    R.x = x;
    R.f = f0;
    R.g = hg;
    Mat V;
    if(!spd_inverse(H, n, V)) {
        make_pos_def(H, n);
        spd_inverse(H, n, V);
    }
    R.V   = V;
    R.edm = 0.5 * dot(hg, matvec(V, hg));
    MnUserParameterState st;
    fill_state(F, R, up, st);
    st.valid_ = true;
    return st;
}

// ------------------------------------------------------------ MnMinos

std::pair<double, double> MnMinos::operator()(unsigned int par, unsigned int maxcalls) const {
    const MnUserParameters &base = min_.UserParameters();
    const double up              = fcn_.Up();
    const double f0              = min_.Fval();
    const double val             = base.Value(par);
    const double err             = base.Error(par) > 0 ? base.Error(par) : 1.0;
    const MinuitParameter &P     = base.Parameters().at(par);
    // profile: minimise every other parameter with `par` fixed at v
    auto profile = [&](double v) {
        MnUserParameters p = base;
        p.SetValue(par, v);
        p.Fix(par);
        MnMigrad m(fcn_, p, strategy_);
        int level         = MnPrint::SetGlobalLevel(0);
        FunctionMinimum r = m(maxcalls, 0.01);
        MnPrint::SetGlobalLevel(level);
        return r.Fval();
    };
    double res[2] = {0, 0};
    for(int side = 0; side < 2; side++) {
        const double sgn = side == 0 ? -1.0 : 1.0;
        // g(t) = sqrt((profile(val + sgn t err) - f0) / Up) - 1 is linear in t for a
        // parabolic profile; secant iterations on t starting from 1 and 1.2
        double t0 = 1.0, t1 = 1.2;
        auto g = [&](double t) {
            double v = val + sgn * t * err;
            if(P.has_lower && v < P.lower)
                v = P.lower;
            if(P.has_upper && v > P.upper)
                v = P.upper;
            double d = (profile(v) - f0) / up;
            return std::sqrt(std::max(d, 0.0)) - 1.0;
        };
        double g0 = g(t0), g1 = g(t1);
        for(int it = 0; it < 8 && std::fabs(g1) > 1e-4; it++) {
            if(g1 == g0)
                break;
            double t2 = t1 - g1 * (t1 - t0) / (g1 - g0);
            t2        = std::max(0.05, std::min(t2, 10.0));
            t0        = t1;
            g0        = g1;
            t1        = t2;
            g1        = g(t1);
        }
        double v = val + sgn * t1 * err;
        if(P.has_lower && v < P.lower)
            v = P.lower;
        if(P.has_upper && v > P.upper)
            v = P.upper;
        res[side] = v - val;
    }
    return {res[0], res[1]};
}

// ------------------------------------------------------------ MnScan

std::vector<std::pair<double, double>> MnScan::Scan(unsigned int par, unsigned int maxsteps, double low, double high) {
    std::vector<std::pair<double, double>> out;
    const MinuitParameter &P = par_.Parameters().at(par);
    if(low == 0 && high == 0) {
        low  = P.value - 2.0 * P.error;
        high = P.value + 2.0 * P.error;
        if(P.has_lower)
            low = std::max(low, P.lower);
        if(P.has_upper)
            high = std::min(high, P.upper);
    }
    if(maxsteps < 2)
        maxsteps = 2;
    std::vector<double> x = par_.Params();
    double best = (*fcn_)(x), bestv = P.value;
    out.emplace_back(P.value, best);
    for(unsigned int s = 0; s < maxsteps; s++) {
        double v = low + (high - low) * s / (maxsteps - 1);
        x[par]   = v;
        double f = (*fcn_)(x);
        out.emplace_back(v, f);
        if(f < best) {
            best  = f;
            bestv = v;
        }
    }
    par_.SetValue(par, bestv);
    return out;
}

// ------------------------------------------------------------ printing

int MnPrint::SetGlobalLevel(int level) {
    int old       = g_print_level;
    g_print_level = level;
    return old;
}
int MnPrint::GlobalLevel() { return g_print_level; }

std::ostream &operator<<(std::ostream &os, const MnUserParameterState &st) {
    const auto &p = st.Parameters().Parameters();
    os << "  # name                        value          error\n";
    for(size_t i = 0; i < p.size(); i++) {
        char line[160];
        snprintf(line, sizeof(line), "  %-3zu %-24s %14.8g %14.8g%s\n", i, p[i].name.c_str(), p[i].value, p[i].error,
                 p[i].fixed ? "  (fixed)" : "");
        os << line;
    }
    return os;
}

std::ostream &operator<<(std::ostream &os, const FunctionMinimum &m) {
    os << "Minuit2 minimum: " << (m.IsValid() ? "valid" : "INVALID") << "\n"
       << "  FCN = " << std::setprecision(12) << m.Fval() << "  Edm = " << m.Edm() << "  NFcn = " << m.NFcn()
       << "\n"
       << "  covariance: " << (m.HasAccurateCovar() ? "accurate" : "approximate")
       << (m.HasMadePosDefCovar() ? ", forced positive definite" : "") << "\n"
       << m.UserState();
    return os;
}

}  // namespace Minuit2
}  // namespace ROOT
