// GooPdf.cpp -- normalisation, NLL and evaluation of a PDF through the GPU engine
// (src/PDFs/GooPdf.cu and src/PDFs/GooPdf.cpp of the reference).
#include <goofit/Error.h>
#include <goofit/PDFs/GooPdf.h>

#include <cmath>

namespace GooFit {

int GooPdf::get_event_size() const {
    int n = (int)observablesList.size();
    return fitControl && fitControl->binnedFit() ? -(n + 2) : n;
}

int GooPdf::metric_code() const { return fitControl ? (int)fitControl->getMetric() : GF_METRIC_NLL; }

double GooPdf::reduce_with_metric() const {
    engine::flush();
    double sum = 0;
    engine::check(gf_reduce_metric(engine::ctx(), metric_code(), &sum, nullptr), "reduce_with_metric");
    return sum;
}

void GooPdf::evaluate_with_metric(std::vector<fptype> &results) const {
    engine::flush();
    results.resize(numEntries);
    engine::check(gf_eval_metric(engine::ctx(), metric_code(), results.data(), (int64_t)numEntries),
                  "evaluate_with_metric");
}

std::vector<fptype> GooPdf::evaluate_with_metric() const {
    std::vector<fptype> r;
    evaluate_with_metric(r);
    return r;
}

int GooPdf::get_bin_grid_size() const {
    size_t total = 1;
    for(const Observable &v : observablesList)
        total *= integrationBins > 0 ? (size_t)integrationBins : v.getNumBins();
    return (int)total;
}

fptype GooPdf::get_bin_grid_volume() const {
    fptype ret = 1.0;
    for(const Observable &v : observablesList)
        ret *= v.getUpperLimit() - v.getLowerLimit();
    return ret;
}

double GooPdf::reduce_with_bins() const {
    engine::flush();
    double sum = 0;
    engine::check(gf_reduce_bins(engine::ctx(), (int)functionIdx, normRanges.data(), (int)observablesList.size(), &sum),
                  "reduce_with_bins");
    return sum;
}

// GooPdf::calculateNLL, src/PDFs/GooPdf.cu:254-273
double GooPdf::calculateNLL() {
    normalize();
    if(host_normalizations.at(normalIdx + 1) <= 0)
        GooFit::abort(__FILE__, __LINE__, getName() + " non-positive normalization", this);
    double ret = reduce_with_metric();
    if(0.0 == ret)
        GooFit::abort(__FILE__, __LINE__, getName() + " zero NLL", this);
    return ret;
}

bool GooPdf::normalisationPositive() const { return host_normalizations.at(normalIdx + 1) > 0; }

// GooPdf::normalize, src/PDFs/GooPdf.cu:340-396
fptype GooPdf::normalize() {
    if(fitControl && !fitControl->metricIsPdf()) {
        setNormalization(1.0);
        return 1.0;
    }
    fptype ret = 1;
    if(hasAnalyticIntegral()) {
        for(const Observable &v : observablesList)
            ret *= integrate(v.getLowerLimit(), v.getUpperLimit());
        setNormalization(1.0 / ret);
        return ret;
    }
    size_t totalBins = get_bin_grid_size();
    ret *= get_bin_grid_volume();
    ret /= totalBins;
    fptype sum = reduce_with_bins();
    if(std::isnan(sum))
        GooFit::abort(__FILE__, __LINE__, getName() + " NaN in normalization", this);
    else if(0 >= sum)
        GooFit::abort(__FILE__, __LINE__, "Non-positive normalization", this);
    ret *= sum;
    if(0 == ret)
        GooFit::abort(__FILE__, __LINE__, "Zero integral");
    setNormalization(1.0 / ret);
    return ret;
}

// A normalisation no floating parameter moves stays at its value in the template
// array; any other numerical normalisation has no device op.
bool GooPdf::batchNormOps(std::vector<gf_norm_op> &, std::vector<double> &) const {
    if(fitControl && !fitControl->metricIsPdf())
        return true;
    for(const Variable &v : getParameters())
        if(!v.IsFixed())
            return false;
    return true;
}

fptype GooPdf::integrate(fptype, fptype) const { return 0; }
bool GooPdf::hasAnalyticIntegral() const { return false; }

std::vector<fptype> GooPdf::evaluateAtPoints(Observable var) {
    setFitControl(std::make_shared<EvalFit>());
    setIndices();
    normalize();
    UnbinnedDataSet tempdata(observablesList);
    double step = var.getBinSize();
    for(size_t i = 0; i < var.getNumBins(); ++i) {
        var.setValue(var.getLowerLimit() + (i + 0.5) * step);
        tempdata.addEvent();
    }
    auto *old = getData();
    setData(&tempdata);
    normalize();
    std::vector<fptype> h = evaluate_with_metric();
    std::vector<fptype> res(var.getNumBins());
    for(size_t i = 0; i < var.getNumBins(); ++i)
        res[i] = h[i] * host_normalizations.at(normalIdx + 1);
    setData(old);
    return res;
}

fptype GooPdf::getValue(EvalFunc evalfunc) {
    if(evalfunc == EvalFunc::Prob)
        setFitControl(std::make_shared<ProbFit>());
    else if(evalfunc == EvalFunc::Eval)
        setFitControl(std::make_shared<EvalFit>());
    else if(evalfunc == EvalFunc::NLL)
        setFitControl(std::make_shared<UnbinnedNllFit>());
    else
        throw GeneralError("That EvalFunc is not supported");
    setIndices();
    normalize();
    UnbinnedDataSet point(observablesList);
    point.addEvent();
    auto *old = getData();
    setData(&point);
    auto results = evaluate_with_metric();
    setData(old);
    return results.at(0);
}

std::vector<fptype> GooPdf::evaluateAtData(EvalFunc evalfunc) {
    auto fc = fitControl;
    if(evalfunc == EvalFunc::Prob)
        setFitControl(std::make_shared<ProbFit>());
    else if(evalfunc == EvalFunc::Eval)
        setFitControl(std::make_shared<EvalFit>());
    else if(evalfunc == EvalFunc::NLL)
        setFitControl(std::make_shared<UnbinnedNllFit>());
    else
        throw GeneralError("That EvalFunc is not supported");
    normalize();
    std::vector<fptype> out = evaluate_with_metric();
    setFitControl(fc);
    return out;
}

std::vector<std::vector<fptype>> GooPdf::getCompProbsAtDataPoints() {
    auto fc = fitControl;
    setFitControl(std::make_shared<ProbFit>());
    normalize();
    std::vector<std::vector<fptype>> values;
    values.push_back(evaluate_with_metric());
    for(auto *comp : components) {
        comp->setIndices();
        comp->normalize();
        values.push_back(evaluate_with_metric());
    }
    setFitControl(fc);
    return values;
}

UnbinnedDataSet GooPdf::makeGrid() {
    UnbinnedDataSet grid{getObservables()};
    grid.fillWithGrid();
    return grid;
}

void GooPdf::initialize() {
    if(!fitControl)
        setFitControl(std::make_shared<UnbinnedNllFit>());
    PdfBase::initializeIndices();
    setMetrics();
}

void GooPdf::scan(Observable var, std::vector<fptype> &values) {
    fptype step = (var.getUpperLimit() - var.getLowerLimit()) / var.getNumBins();
    values.clear();
    for(fptype v = var.getLowerLimit() + 0.5 * step; v < var.getUpperLimit(); v += step) {
        var.setValue(v);
        copyParams();
        values.push_back(calculateNLL());
    }
}

void GooPdf::setParameterConstantness(bool constant) {
    std::vector<Variable> pars = getParameters();
    for(Variable &p : pars)
        p.setFixed(constant);
}

void GooPdf::setFitControl(std::shared_ptr<FitControl> fc) {
    for(auto *c : components)
        c->setFitControl(fc);
    fitControl = fc;
    setMetrics();
    setIndices();
}

void GooPdf::setIndices() {
    if(!fitControl)
        fitControl = std::make_shared<UnbinnedNllFit>();
    setMetrics();
    PdfBase::setIndices();
    pre_run();
}

}  // namespace GooFit
