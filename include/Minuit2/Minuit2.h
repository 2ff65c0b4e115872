// Minuit2-compatible minimiser (MIGRAD, HESSE, MINOS, scan) -- the API surface
// GooFit's fitting layer binds (src/goofit/FitManagerMinuit2.cpp:31-124,
// src/goofit/Params.cpp:14-57, include/goofit/fitting/FCN.h:11-32).  Minuit2 is
// an external dependency of the reference (extern/Minuit2, GooFit/Minuit2) that
// is not available here, so this is a self-contained implementation of the
// published variable-metric method (F. James, MINUIT; Davidon-Fletcher-Powell /
// BFGS update with parabolic line search, numerical gradients and a final
// numerical Hessian) behind the same class and method names.
#pragma once

#include <cstddef>
#include <iosfwd>
#include <string>
#include <utility>
#include <vector>

namespace ROOT {
namespace Minuit2 {

/// Function to minimise (Minuit2 FCNBase).
class FCNBase {
  public:
    virtual ~FCNBase() = default;
    virtual double operator()(const std::vector<double> &x) const = 0;
    /// Change of the function value that defines one standard deviation (0.5 for -log L).
    virtual double Up() const = 0;
    virtual double ErrorDef() const { return Up(); }
    /// The function at several independent points, in order.  No counterpart in
    /// Minuit2: the minimiser hands over the points of a derivative step together
    /// (the 2n points of one gradient cycle, the points of a Hessian) so that a
    /// function evaluated on a GPU can serve them in one launch.  The default
    /// evaluates them one by one; an override must return what operator() would.
    virtual std::vector<double> EvaluateMany(const std::vector<std::vector<double>> &xs) const {
        std::vector<double> out;
        out.reserve(xs.size());
        for(const auto &x : xs)
            out.push_back((*this)(x));
        return out;
    }
};

/// One user parameter: value, step/error, optional limits, fixed flag.
struct MinuitParameter {
    std::string name;
    double value = 0, error = 0, lower = 0, upper = 0;
    bool has_lower = false, has_upper = false, fixed = false;
};

/// Ordered set of named parameters (Minuit2 MnUserParameters).
class MnUserParameters {
  public:
    MnUserParameters() = default;
    virtual ~MnUserParameters() = default;

    /// Constant parameter.
    bool Add(const std::string &name, double val);
    /// Free parameter without limits.
    bool Add(const std::string &name, double val, double err);
    /// Free parameter with limits [low, up].
    bool Add(const std::string &name, double val, double err, double low, double up);

    unsigned int Index(const std::string &name) const;
    const std::string &GetName(unsigned int i) const { return pars_.at(i).name; }
    void SetValue(unsigned int i, double v) { pars_.at(i).value = v; }
    void SetError(unsigned int i, double e) { pars_.at(i).error = e; }
    void SetLimits(unsigned int i, double low, double up);
    void SetLowerLimit(unsigned int i, double low);
    void SetUpperLimit(unsigned int i, double up);
    void RemoveLimits(unsigned int i);
    void Fix(unsigned int i) { pars_.at(i).fixed = true; }
    void Release(unsigned int i) { pars_.at(i).fixed = false; }
    double Value(unsigned int i) const { return pars_.at(i).value; }
    double Error(unsigned int i) const { return pars_.at(i).error; }
    unsigned int NParameters() const { return (unsigned int)pars_.size(); }
    unsigned int VariableParameters() const;
    const std::vector<MinuitParameter> &Parameters() const { return pars_; }
    std::vector<double> Params() const;
    std::vector<double> Errors() const;

  protected:
    std::vector<MinuitParameter> pars_;
};

/// Parameter values, errors and external covariance after a minimisation.
class MnUserParameterState {
  public:
    MnUserParameterState() = default;
    explicit MnUserParameterState(const MnUserParameters &p)
        : pars_(p) {}

    double Value(unsigned int i) const { return pars_.Value(i); }
    double Error(unsigned int i) const { return pars_.Error(i); }
    double Value(const std::string &name) const { return pars_.Value(pars_.Index(name)); }
    double Error(const std::string &name) const { return pars_.Error(pars_.Index(name)); }
    const MnUserParameters &Parameters() const { return pars_; }
    MnUserParameters &Parameters() { return pars_; }
    /// Covariance between external parameters i and j (0 for fixed parameters).
    double Covariance(unsigned int i, unsigned int j) const;
    bool HasCovariance() const { return !cov_.empty(); }
    double Fval() const { return fval_; }
    double Edm() const { return edm_; }
    unsigned int NFcn() const { return nfcn_; }
    bool IsValid() const { return valid_; }

    // filled by the minimiser
    MnUserParameters pars_;
    std::vector<double> cov_;  // n x n over all external parameters
    double fval_ = 0, edm_ = 0;
    unsigned int nfcn_ = 0;
    bool valid_        = false;
};

/// Strategy level 0 (fast), 1 (default: Hessian at the end), 2 (careful).
class MnStrategy {
  public:
    explicit MnStrategy(unsigned int level = 1)
        : level_(level) {}
    unsigned int Strategy() const { return level_; }
    unsigned int GradientNCycles() const { return level_ == 0 ? 2 : (level_ == 1 ? 3 : 5); }
    unsigned int HessianNCycles() const { return level_ == 0 ? 3 : (level_ == 1 ? 5 : 7); }

  private:
    unsigned int level_;
};

/// Covariance of the variable (non-fixed) parameters, row-major (Minuit2 MnUserCovariance).
class MnUserCovariance {
  public:
    MnUserCovariance() = default;
    MnUserCovariance(unsigned int n, std::vector<double> data)
        : n_(n)
        , data_(std::move(data)) {}
    unsigned int Nrow() const { return n_; }
    unsigned int size() const { return (unsigned int)data_.size(); }
    double operator()(unsigned int i, unsigned int j) const { return data_.at((size_t)i * n_ + j); }
    const std::vector<double> &Data() const { return data_; }

  private:
    unsigned int n_ = 0;
    std::vector<double> data_;
};

/// Result of a minimisation (Minuit2 FunctionMinimum).
class FunctionMinimum {
  public:
    bool IsValid() const { return valid_; }
    bool HasCovariance() const { return has_cov_; }
    bool HasValidCovariance() const { return has_cov_ && pos_def_; }
    bool HasValidParameters() const { return valid_params_; }
    bool HasAccurateCovar() const { return accurate_; }
    bool HasPosDefCovar() const { return pos_def_; }
    bool HasMadePosDefCovar() const { return made_pos_def_; }
    bool HesseFailed() const { return hesse_failed_; }
    bool IsAboveMaxEdm() const { return above_max_edm_; }
    bool HasReachedCallLimit() const { return call_limit_; }
    double Fval() const { return state_.fval_; }
    double Edm() const { return state_.edm_; }
    unsigned int NFcn() const { return state_.nfcn_; }
    double Up() const { return up_; }
    const MnUserParameterState &UserState() const { return state_; }
    const MnUserParameters &UserParameters() const { return state_.pars_; }
    /// Covariance over the variable parameters (in their order of appearance).
    MnUserCovariance UserCovariance() const {
        std::vector<unsigned int> var;
        for(unsigned int i = 0; i < state_.pars_.NParameters(); i++)
            if(!state_.pars_.Parameters()[i].fixed)
                var.push_back(i);
        std::vector<double> c(var.size() * var.size(), 0.0);
        if(state_.HasCovariance())
            for(size_t a = 0; a < var.size(); a++)
                for(size_t b = 0; b < var.size(); b++)
                    c[a * var.size() + b] = state_.Covariance(var[a], var[b]);
        return MnUserCovariance((unsigned int)var.size(), c);
    }

    // filled by the minimiser
    MnUserParameterState state_;
    double up_         = 0.5;
    bool valid_        = false;
    bool valid_params_ = false;
    bool has_cov_      = false;
    bool accurate_     = false;
    bool pos_def_      = false;
    bool made_pos_def_ = false;
    bool hesse_failed_ = false;
    bool above_max_edm_ = false;
    bool call_limit_   = false;
    // internal-coordinate state, used by MnMinos to restart from the minimum
    std::vector<double> x_int_, v_int_;
};

std::ostream &operator<<(std::ostream &os, const FunctionMinimum &min);
std::ostream &operator<<(std::ostream &os, const MnUserParameterState &st);

/// Variable-metric minimisation (MIGRAD).
class MnMigrad {
  public:
    MnMigrad(const FCNBase &fcn, const MnUserParameters &par, const MnStrategy &str = MnStrategy(1))
        : fcn_(fcn)
        , par_(par)
        , strategy_(str) {}
    MnMigrad(const FCNBase &fcn, const MnUserParameters &par, unsigned int strategy)
        : fcn_(fcn)
        , par_(par)
        , strategy_(strategy) {}

    /// maxfcn 0 = 200 + 100 n + 5 n^2; the EDM target is 0.002 * tolerance * Up.
    FunctionMinimum operator()(unsigned int maxfcn = 0, double tolerance = 0.1);

  private:
    const FCNBase &fcn_;
    MnUserParameters par_;
    MnStrategy strategy_;
};

/// Full numerical second-derivative matrix at a point (MnHesse).
class MnHesse {
  public:
    explicit MnHesse(const MnStrategy &str = MnStrategy(1))
        : strategy_(str) {}
    MnUserParameterState operator()(const FCNBase &fcn, const MnUserParameters &par, unsigned int maxcalls = 0) const;

  private:
    MnStrategy strategy_;
};

/// Asymmetric errors by profiling the function (MINOS).
class MnMinos {
  public:
    MnMinos(const FCNBase &fcn, const FunctionMinimum &min, const MnStrategy &str = MnStrategy(1))
        : fcn_(fcn)
        , min_(min)
        , strategy_(str) {}
    MnMinos(const FCNBase &fcn, const FunctionMinimum &min, unsigned int strategy)
        : fcn_(fcn)
        , min_(min)
        , strategy_(strategy) {}
    /// (lower, upper) error of parameter par: the values where min over the other
    /// parameters of the function rises by Up.
    std::pair<double, double> operator()(unsigned int par, unsigned int maxcalls = 0) const;
    double Lower(unsigned int par, unsigned int maxcalls = 0) const { return (*this)(par, maxcalls).first; }
    double Upper(unsigned int par, unsigned int maxcalls = 0) const { return (*this)(par, maxcalls).second; }

  private:
    const FCNBase &fcn_;
    const FunctionMinimum &min_;
    MnStrategy strategy_;
};

/// One-dimensional parameter scans (MnScan).
class MnScan {
  public:
    MnScan(const FCNBase &fcn, const MnUserParameters &par, unsigned int strategy = 1)
        : fcn_(&fcn)
        , par_(par)
        , strategy_(strategy) {}
    /// (value, function) pairs over [low, high] (default: value +- 2 errors, or the limits).
    std::vector<std::pair<double, double>> Scan(unsigned int par, unsigned int maxsteps = 41, double low = 0.,
                                                double high = 0.);

  private:
    const FCNBase *fcn_;
    MnUserParameters par_;
    unsigned int strategy_;
};

/// Print level shared by the minimiser (0 quiet .. 3 verbose).
class MnPrint {
  public:
    static int SetGlobalLevel(int level);
    static int GlobalLevel();
    static void SetLevel(int level) { SetGlobalLevel(level); }
    static int Level() { return GlobalLevel(); }
};

}  // namespace Minuit2
}  // namespace ROOT
