// PdfBase.cpp -- tree registration, flattening and data upload
// (src/goofit/PdfBase.cpp and src/PDFs/PdfBase.cu of the reference).
#include <goofit/BinnedDataSet.h>
#include <goofit/Error.h>
#include <goofit/FitControl.h>
#include <goofit/FitManager.h>
#include <goofit/PDFs/GooPdf.h>
#include <goofit/PdfBase.h>
#include <goofit/UnbinnedDataSet.h>

#include <algorithm>
#include <iostream>
#include <memory>
#include <numeric>
#include <random>

namespace GooFit {

namespace {
template <typename T>
bool contains(const std::vector<T> &list, const T &item) {
    return std::find(list.begin(), list.end(), item) != list.end();
}
}  // namespace

[[noreturn]] void abort(std::string file, int line, std::string reason, const PdfBase *pdf) {
    std::cout << "GooFit is aborting at " << file << ":" << line << ": " << reason << "\n";
    if(pdf) {
        std::vector<Variable> pars = pdf->getParameters();
        std::cout << "Parameters of " << pdf->getName() << ":\n";
        for(const Variable &v : pars)
            std::cout << "  " << v.getName() << " = " << v.getValue() << "\n";
    }
    throw GeneralError(reason);
}

void PdfBase::pre_run() { engine::mark_program_dirty(); }

void PdfBase::pre_call() {
    // parameter values are compared and uploaded by engine::flush() before the next device call
}

void PdfBase::copyParams() { updateParameters(); }

// Placeholder blocks so that parametersIdx etc. are valid before the first setIndices()
// (PdfBase::initializeIndices, src/PDFs/PdfBase.cu:45-95).
void PdfBase::initializeIndices() {
    parametersIdx = (unsigned int)host_parameters.size();
    host_parameters.push_back((fptype)parametersList.size());
    host_parameter_name.push_back("nParam");
    for(auto &p : parametersList) {
        host_parameters.push_back(p.getValue());
        host_parameter_name.push_back(p.getName());
    }
    constantsIdx = (unsigned int)host_constants.size();
    host_constants.push_back((fptype)constantsList.size());
    for(fptype c : constantsList)
        host_constants.push_back(c);
    observablesIdx = (unsigned int)host_observables.size();
    host_observables.push_back((fptype)observablesList.size());
    for(auto &o : observablesList)
        host_observables.push_back(o.getValue());
    normalIdx = (unsigned int)host_normalizations.size();
    host_normalizations.push_back(1.0);
    host_normalizations.push_back(0.0);
}

void PdfBase::registerFunction(std::string name, int function_code) {
    reflex_name_   = std::move(name);
    function_code_ = function_code;
}

void PdfBase::recursiveSetIndices() {
    if(reflex_name_.empty() || function_code_ == GF_FN_NONE)
        throw GeneralError("A PDF must either provide a function name and function code or override "
                           "recursiveSetIndices\nCalled by {}",
                           getName());
    functionIdx = (unsigned int)host_function_table.size();
    host_function_table.push_back(function_code_);
    host_function_name.push_back(getName());
    populateArrays();
}

// Blocks [count, values...] of this node, then of every component, depth first
// (PdfBase::populateArrays, src/PDFs/PdfBase.cu:268-311).
void PdfBase::populateArrays() {
    parametersIdx = (unsigned int)host_parameters.size();
    host_parameters.push_back((fptype)parametersList.size());
    host_parameter_name.push_back("nParam");
    for(auto &p : parametersList) {
        host_parameters.push_back(p.getValue());
        host_parameter_name.push_back(p.getName());
    }
    constantsIdx = (unsigned int)host_constants.size();
    host_constants.push_back((fptype)constantsList.size());
    for(fptype c : constantsList)
        host_constants.push_back(c);
    observablesIdx = (unsigned int)host_observables.size();
    host_observables.push_back((fptype)observablesList.size());
    for(auto &o : observablesList)
        host_observables.push_back((fptype)o.getIndex());
    normalIdx = (unsigned int)host_normalizations.size();
    host_normalizations.push_back(1.0);
    host_normalizations.push_back(cachedNormalization);
    for(auto *c : components)
        c->recursiveSetIndices();
    generateNormRange();
}

void PdfBase::setIndices() {
    engine::bump_layout_epoch();
    host_parameters.clear();
    host_constants.clear();
    host_observables.clear();
    host_normalizations.clear();
    host_function_table.clear();
    host_function_name.clear();
    host_parameter_name.clear();
    recursiveSetIndices();
}

// PdfBase::updateVariable, src/PDFs/PdfBase.cu:121-129: every parameter of the
// tree with var's name takes the value.  Copies of one Variable share their value,
// so the storage test settles most entries without comparing names.
void PdfBase::updateVariable(const Variable &var, fptype newValue) {
    for(auto &p : parametersList)
        if(p == var || p.getName() == var.getName())
            p.setValue(newValue);
    for(auto *c : components)
        c->updateVariable(var, newValue);
}

void PdfBase::updateParameters() {
    for(size_t i = 0; i < parametersList.size(); i++)
        if(parametersIdx + i + 1 < host_parameters.size())
            host_parameters[parametersIdx + i + 1] = parametersList[i].getValue();
    for(auto *c : components)
        c->updateParameters();
    pre_call();
}

void PdfBase::matchingParameters(const Variable &var, std::vector<Variable> &out) const {
    for(const auto &p : parametersList)
        if(p == var || p.getName() == var.getName())
            out.push_back(p);
    for(auto *c : components)
        c->matchingParameters(var, out);
}

void PdfBase::parameterSlots(std::vector<std::pair<Variable, unsigned int>> &out) const {
    for(size_t i = 0; i < parametersList.size(); i++)
        if(parametersIdx + i + 1 < host_parameters.size())
            out.emplace_back(parametersList[i], (unsigned int)(parametersIdx + i + 1));
    for(auto *c : components)
        c->parameterSlots(out);
}

void PdfBase::recursiveSetNormalization(fptype norm, bool subpdf) {
    host_normalizations.at(normalIdx + 1) = norm;
    cachedNormalization                   = norm;
    for(auto *c : components)
        c->recursiveSetNormalization(norm, true);
}

void PdfBase::setNormalization(const fptype &v) {
    cachedNormalization = v;
    if(normalIdx + 1 < host_normalizations.size())
        host_normalizations[normalIdx + 1] = v;
}

void PdfBase::registerParameter(Variable var) { parametersList.push_back(var); }
void PdfBase::registerConstant(fptype value) { constantsList.push_back(value); }

void PdfBase::unregisterParameter(Variable var) {
    auto pos = std::find(parametersList.begin(), parametersList.end(), var);
    if(pos != parametersList.end())
        parametersList.erase(pos);
    for(auto *c : components)
        c->unregisterParameter(var);
}

std::vector<Variable> PdfBase::getParameters() const {
    std::vector<Variable> ret(parametersList.begin(), parametersList.end());
    for(const PdfBase *c : components)
        for(const Variable &v : c->getParameters())
            if(!contains(ret, v))
                ret.push_back(v);
    return ret;
}

Variable *PdfBase::getParameterByName(std::string n) {
    for(Variable &p : parametersList)
        if(p.getName() == n)
            return &p;
    for(auto *c : components) {
        Variable *v = c->getParameterByName(n);
        if(v)
            return v;
    }
    return nullptr;
}

std::vector<Observable> PdfBase::getObservables() const {
    std::vector<Observable> ret(observablesList.begin(), observablesList.end());
    for(const PdfBase *c : components)
        for(const Observable &o : c->getObservables())
            if(!contains(ret, o))
                ret.push_back(o);
    return ret;
}

void PdfBase::registerObservable(Observable obs) {
    if(contains(observablesList, obs))
        return;
    observablesList.push_back(obs);
}

void PdfBase::setupObservables() {
    int counter = 0;
    for(auto &o : observablesList)
        o.setIndex(counter++);
}

void PdfBase::setIntegrationFineness(int i) {
    integrationBins = i;
    generateNormRange();
}

bool PdfBase::parametersChanged() const {
    return std::any_of(parametersList.begin(), parametersList.end(), [](const Variable &v) { return v.getChanged(); });
}

// [lower, upper, bins] of each of this PDF's observables (PdfBase::generateNormRange,
// src/PDFs/PdfBase.cu:532-557).
void PdfBase::generateNormRange() {
    normRanges.clear();
    for(const Observable &v : observablesList) {
        normRanges.push_back(v.getLowerLimit());
        normRanges.push_back(v.getUpperLimit());
        normRanges.push_back(integrationBins > 0 ? (fptype)integrationBins : (fptype)v.getNumBins());
    }
}

// Number the observables, flatten the tree and upload the events as SoA columns
// (PdfBase::setData, src/PDFs/PdfBase.cu:329-437; the reference packs an AoS array).
void PdfBase::setData(DataSet *data) {
    setupObservables();
    setIndices();
    data_ = data;
    if(data == nullptr)
        return;
    const size_t dims = observablesList.size();
    // GF_MAX_OBS = 8 event columns: a binned row adds the content and the volume/error
    const bool is_binned = dynamic_cast<BinnedDataSet *>(data) != nullptr;
    if(dims + (is_binned ? 2 : 0) > 8)
        throw GeneralError("{}: {} observables (the engine handles at most 6 with bins, 8 without)", getName(), dims);
    gf_ctx *c = engine::ctx();
    if(auto *unbinned = dynamic_cast<UnbinnedDataSet *>(data)) {
        numEntries = (unsigned int)data->getNumEvents();
        numEvents  = numEntries;
        std::vector<const double *> cols(dims, nullptr);
        for(const Observable &v : observablesList) {
            const std::vector<fptype> &col = unbinned->getColumn(v);
            cols.at(v.getIndex())          = col.data();
        }
        engine::check(gf_events_upload(c, (int)dims, (int64_t)numEntries, cols.data()), "gf_events_upload");
        for(const Observable &v : observablesList)
            if(v.isEventNumber() && numEntries > 0) {
                if(gf_events_set_identity_column(c, v.getIndex()) != GF_OK) {
                    // the event numbers are not the row index: the column is read as is
                }
            }
        return;
    }
    auto *binned = dynamic_cast<BinnedDataSet *>(data);
    if(!binned)
        throw GeneralError("Dataset must be binned or unbinned!");
    // PdfBase.cu:439-527: one row per bin: the bin centre of each observable at its
    // index, then the content, then the volume (or the error for BinnedErrorFit);
    // c_totalEvents is the number of bins.
    numEvents  = 0;
    numEntries = (unsigned int)binned->getNumBins();
    if(!fitControl || !fitControl->binnedFit())
        setFitControl(std::make_shared<BinnedNllFit>());
    std::vector<std::vector<double>> cols(dims + 2, std::vector<double>(numEntries));
    for(unsigned int i = 0; i < numEntries; ++i) {
        for(const Observable &v : observablesList)
            cols[v.getIndex()][i] = binned->getBinCenter(v, i);
        cols[dims][i]     = binned->getBinContent(i);
        cols[dims + 1][i] = fitControl->binErrors() ? binned->getBinError(i) : binned->getBinVolume(i);
        numEvents += binned->getBinContent(i);
    }
    std::vector<const double *> ptrs;
    for(auto &col : cols)
        ptrs.push_back(col.data());
    engine::check(gf_events_upload(c, (int)(dims + 2), (int64_t)numEntries, ptrs.data()), "gf_events_upload");
}

// PdfBase::fillMCDataSimple, src/goofit/PdfBase.cpp:154-203.  Events of the lone
// observable drawn from the PDF tabulated at its bin centres: a cell by inverse CDF
// of the tabulated values, then a uniform position inside the cell.  Each event takes
// two draws of a std::mt19937 seeded with `seed` (0: a random seed) through
// std::uniform_real_distribution -- the cell's uniform in [0, 1), then the offset in
// [-1/2, 1/2) bin widths -- so a given seed yields the reference's events.  The new
// events are appended to the PDF's current data set, which is set again at the end
// so that the device holds them.
void PdfBase::fillMCDataSimple(size_t events, unsigned int seed) {
    if(observablesList.size() != 1)
        throw GeneralError("You can only fill MC on a 1D dataset with a simple fill");
    DataSet *target = getData();
    if(target == nullptr)
        throw GeneralError("Can't run on a PDF with no DataSet to fill!");
    Observable obs = observablesList.front();

    // the PDF at the bin centres (the grid kernel's cells), evaluated on the device
    UnbinnedDataSet cells{obs};
    cells.fillWithGrid();
    setData(&cells);
    std::vector<double> cdf = dynamic_cast<GooPdf &>(*this).getCompProbsAtDataPoints().front();
    // running sum in cell order, as a fraction of the total
    for(size_t k = 1; k < cdf.size(); k++)
        cdf[k] += cdf[k - 1];
    const double total = cdf.back();
    for(double &c : cdf)
        c /= total;

    std::mt19937 engine(seed != 0 ? seed : std::random_device{}());
    std::uniform_real_distribution<> pick_cell(0.0, 1.0), within_cell(-.5, .5);
    const double width = obs.getBinSize();
    for(size_t i = 0; i < events; i++) {
        const double u    = pick_cell(engine);
        const size_t cell = (size_t)(std::lower_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
        obs.setValue(cells.getValue(obs, cell) + width * within_cell(engine));
        target->addEvent();
    }
    setData(target);
}

ROOT::Minuit2::FunctionMinimum PdfBase::fitTo(DataSet *data, int verbosity) {
    auto *old = getData();
    setData(data);
    auto m = fit(verbosity);
    setData(old);
    return m;
}

ROOT::Minuit2::FunctionMinimum PdfBase::fit(int verbosity) {
    FitManager fitter{this};
    fitter.setVerbosity(verbosity);
    return fitter.fit();
}

bool PdfBase::areParamsandConstantsEqualByVal(const PdfBase &other) const {
    auto a = getParameters(), b = other.getParameters();
    if(a.size() != b.size())
        return false;
    for(size_t i = 0; i < a.size(); i++)
        if(!a[i].isEqualNameValLimits(b[i]))
            return false;
    return constantsList == other.constantsList;
}

std::ostream &operator<<(std::ostream &out, const PdfBase &pdf) {
    out << "GooPdf::" << pdf.getPdfName() << "(\"" << pdf.getName() << "\") :\n";
    out << "  Device function: " << pdf.reflex_name_ << "\n";
    out << "  Observable" << (pdf.observablesList.size() > 1 ? "s" : "") << ": ";
    for(size_t i = 0; i < pdf.observablesList.size(); i++)
        out << pdf.observablesList[i].getName() << (i + 1 == pdf.observablesList.size() ? "" : ", ");
    out << "\n  Parameter" << (pdf.parametersList.size() > 1 ? "s" : "") << ": ";
    for(size_t i = 0; i < pdf.parametersList.size(); i++)
        out << pdf.parametersList[i].getName() << (i + 1 == pdf.parametersList.size() ? "" : ", ");
    out << "\n";
    return out;
}

}  // namespace GooFit
