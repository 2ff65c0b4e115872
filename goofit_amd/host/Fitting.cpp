// Fitting.cpp -- FCN, Params and FitManagerMinuit2 (src/goofit/{FCN,Params,
// FitManagerMinuit2}.cpp of the reference) on the Minuit2-compatible minimiser.
#include <goofit/Error.h>
#include <goofit/Engine.h>
#include <goofit/PDFs/GooPdf.h>
#include <goofit/PdfBase.h>
#include <goofit/fitting/FCN.h>
#include <goofit/fitting/FitManagerMinuit2.h>
#include <goofit/fitting/Params.h>

#include <Minuit2/FunctionMinimum.h>
#include <Minuit2/MnMigrad.h>
#include <Minuit2/MnMinos.h>
#include <Minuit2/MnPrint.h>

#include <chrono>
#include <iostream>

namespace GooFit {

// ---------------------------------------------------------------- Params

// The Minuit side of each free or fixed Variable of the tree.  A Variable enters
// Minuit2 as a constant when fixed, as an unbounded parameter when its limits
// coincide, and as a bounded one otherwise (the three MnUserParameters::Add forms
// the reference uses, src/goofit/Params.cpp:14-35); its Minuit index is stored in
// the Variable (shared by every copy of it in the tree).
namespace {
unsigned int enter_minuit(Minuit2::MnUserParameters &up, Variable &var) {
    const std::string &name = var.getName();
    bool fresh;
    if(var.IsFixed())
        fresh = up.Add(name, var.getValue());
    else if(var.getLowerLimit() == var.getUpperLimit())
        fresh = up.Add(name, var.getValue(), var.getError());
    else
        fresh = up.Add(name, var.getValue(), var.getError(), var.getLowerLimit(), var.getUpperLimit());
    // two distinct parameters with one name would share one Minuit slot
    if(!fresh)
        throw std::runtime_error("Params: the parameter name " + name + " is used by two parameters of the tree");
    const unsigned int k = up.Index(name);
    var.setFitterIndex((int)k);
    return k;
}
}  // namespace

Params::Params(PdfBase &pdf)
    : vars_(pdf.getParameters())
    , pdf_(&pdf) {
    for(Variable &var : vars_)
        enter_minuit(*this, var);
    num_ = vars_.size();
}

// Minuit state -> Variables and this parameter set, index by index.
void Params::SetGooFitParams(const Minuit2::MnUserParameterState &input) {
    for(Variable &var : vars_) {
        const unsigned int k = (unsigned int)var.getFitterIndex();
        const double value = input.Value(k), error = input.Error(k);
        var.setValue(value);
        var.setError(error);
        SetValue(k, value);
        SetError(k, error);
    }
}

// The Minuit vector of the Variables' current values (slot = Minuit index).
std::vector<double> Params::make_minuit_vector() const {
    std::vector<double> x((size_t)max_fitter_index(vars_) + 1, 0.0);
    for(const Variable &var : vars_)
        x.at((size_t)var.getFitterIndex()) = var.getValue();
    return x;
}

// Params::from_minuit_vector, src/goofit/Params.cpp:47-57: the changed flags, the
// values minus the blinding offsets into every parameter updateVariable would set,
// then the parameter array as updateParameters fills it -- the tree walks of those
// two calls planned once per layout and replayed as flat loops.
void Params::from_minuit_vector(const std::vector<double> &values, bool force_changed) {
    if(plan_epoch_ != engine::layout_epoch() || plan_nhost_ != host_parameters.size()) {
        targets_.assign(vars_.size(), {});
        for(size_t k = 0; k < vars_.size(); k++)
            pdf_->matchingParameters(vars_[k], targets_[k]);
        slots_.clear();
        pdf_->parameterSlots(slots_);
        plan_epoch_ = engine::layout_epoch();
        plan_nhost_ = host_parameters.size();
    }
    for(size_t k = 0; k < vars_.size(); k++) {
        Variable &var  = vars_[k];
        const double v = values.at(var.getFitterIndex());
        var.setChanged(force_changed ? true : var.getValue() != v);
        const double nv = v - var.getBlind(Variable::Key());
        for(Variable &t : targets_[k])
            t.setValue(nv);
    }
    if(do_record_)
        recorded_.push_back(values);
    for(const auto &sl : slots_)
        host_parameters[sl.second] = sl.first.getValue();
}

// ---------------------------------------------------------------- FCN

FCN::FCN(Params &params)
    : params_(&params) {
    host_callnumber = 0;
    for(Variable &var : params_->vars_)
        var.setChanged(true);
}

double FCN::operator()(const std::vector<double> &pars) const {
    params_->from_minuit_vector(pars);
    double nll = params_->pdf_->calculateNLL();
    host_callnumber++;
    return nll;
}

double FCN::operator()() const {
    std::vector<double> pars = params_->make_minuit_vector();
    params_->from_minuit_vector(pars, true);
    double nll = params_->pdf_->calculateNLL();
    host_callnumber++;
    return nll;
}

Params *FCN::GetParams() { return params_; }

std::vector<double> FCN::EvaluateMany(const std::vector<std::vector<double>> &xs) const {
    auto *pdf = dynamic_cast<GooPdf *>(params_->pdf_);
    if(!batched_ || xs.size() < 2 || !pdf)
        return Minuit2::FCNBase::EvaluateMany(xs);
    const size_t npts = xs.size();
    std::vector<double> P, N, extra(npts);
    size_t np = 0, nn = 0;
    unsigned long long epoch = 0;
    for(size_t k = 0; k < npts; k++) {
        // what operator() does up to the reduction: set the point, normalise, check
        params_->from_minuit_vector(xs[k]);
        pdf->normalize();
        if(k == 0) {
            epoch = engine::cache_epoch();
        } else if(engine::cache_epoch() != epoch) {
            // this point recomputed a wave cache: the points do not share the caches
            std::vector<double> out(npts);
            for(size_t j = 0; j < npts; j++)
                out[j] = (*this)(xs[j]);
            return out;
        }
        if(!pdf->normalisationPositive())
            GooFit::abort(__FILE__, __LINE__, pdf->getName() + " non-positive normalization", pdf);
        if(k == 0) {
            np = host_parameters.size();
            nn = host_normalizations.size();
            P.reserve(np * npts);
            N.reserve(nn * npts);
        }
        P.insert(P.end(), host_parameters.begin(), host_parameters.end());
        N.insert(N.end(), host_normalizations.begin(), host_normalizations.end());
        extra[k] = pdf->nllExtraTerm();
    }
    engine::flush();
    std::vector<double> out(npts);
    engine::check(gf_reduce_metric_multi(engine::ctx(), pdf->metricCode(), (int)npts, P.data(), (int)np, N.data(),
                                         (int)nn, out.data()),
                  "gf_reduce_metric_multi");
    for(size_t k = 0; k < npts; k++) {
        if(0.0 == out[k])
            GooFit::abort(__FILE__, __LINE__, pdf->getName() + " zero NLL", pdf);
        out[k] += extra[k];
    }
    host_callnumber += (unsigned int)npts;
    return out;
}

// ---------------------------------------------------------------- FitManagerMinuit2

FitManagerMinuit2::FitManagerMinuit2(PdfBase *dat)
    : upar_(*dat)
    , fcn_(upar_) {}

// FitManagerMinuit2::fit, src/goofit/FitManagerMinuit2.cpp:31-119
Minuit2::FunctionMinimum FitManagerMinuit2::fit() {
    int old_level   = Minuit2::MnPrint::SetGlobalLevel(verbosity > 1 ? verbosity - 1 : 0);
    host_callnumber = 0;
    auto t0         = std::chrono::steady_clock::now();
    Minuit2::MnStrategy strat((unsigned int)m_strategy);
    Minuit2::MnMigrad migrad{fcn_, upar_, strat};
    Minuit2::FunctionMinimum min = migrad(maxfcn_, tolerance_);
    if(minos) {
        Minuit2::MnMinos mns{fcn_, min, strat};
        for(Variable &var : upar_.GetGooFitParams()) {
            if(var.IsFixed())
                continue;
            minos_errors.push_back(mns((unsigned int)var.getFitterIndex()));
        }
        if(verbosity > 0) {
            std::cout << "1-sigma minos errors: " << std::endl;
            size_t k = 0;
            for(Variable &var : upar_.GetGooFitParams()) {
                if(var.IsFixed())
                    continue;
                std::cout << var.getName() << ": " << minos_errors[k].first << " " << minos_errors[k].second
                          << std::endl;
                k++;
            }
        }
    }
    last_time_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if(verbosity > 0) {
        std::cout << min;
        std::cout << "The minimization took: " << last_time_ * 1e3 << " ms\n";
        if(min.NFcn() > 0)
            std::cout << "Average time per call: " << last_time_ * 1e3 / min.NFcn() << " ms\n";
    }
    if(min.IsValid() && min.HasCovariance() && !min.IsAboveMaxEdm() && !min.HasReachedCallLimit()) {
        retval_ = FitErrors::Valid;
    } else {
        if(verbosity > 0) {
            std::cout << "HesseFailed: " << min.HesseFailed() << std::endl;
            std::cout << "HasCovariance: " << min.HasCovariance() << std::endl;
            std::cout << "HasValidCovariance: " << min.HasValidCovariance() << std::endl;
            std::cout << "HasValidParameters: " << min.HasValidParameters() << std::endl;
            std::cout << "IsAboveMaxEdm: " << min.IsAboveMaxEdm() << std::endl;
            std::cout << "HasReachedCallLimit: " << min.HasReachedCallLimit() << std::endl;
            std::cout << "HasAccurateCovar: " << min.HasAccurateCovar() << std::endl;
            std::cout << "HasPosDefCovar : " << min.HasPosDefCovar() << std::endl;
            std::cout << "HasMadePosDefCovar : " << min.HasMadePosDefCovar() << std::endl;
        }
        retval_ = FitErrors::InValid;
    }
    upar_.SetGooFitParams(min.UserState());
    Minuit2::MnPrint::SetGlobalLevel(old_level);
    return min;
}

Minuit2::MnScan FitManagerMinuit2::getMnScan() { return Minuit2::MnScan{fcn_, upar_}; }

}  // namespace GooFit
