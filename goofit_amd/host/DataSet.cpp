// DataSet.cpp -- Variable printing, DataSet and UnbinnedDataSet
// (src/goofit/{Variable,DataSet,UnbinnedDataSet}.cpp of the reference).
#include <goofit/Error.h>
#include <goofit/UnbinnedDataSet.h>
#include <goofit/Variable.h>

#include <algorithm>
#include <iostream>

namespace GooFit {

// ---------------------------------------------------------------- Variable

int max_index(const std::vector<Variable> &vars) {
    int m = -1;
    for(const auto &v : vars)
        m = std::max(m, v.getIndex());
    return m;
}
int max_index(const std::vector<Observable> &vars) {
    int m = -1;
    for(const auto &v : vars)
        m = std::max(m, v.getIndex());
    return m;
}
int max_fitter_index(const std::vector<Variable> &vars) {
    int m = -1;
    for(const auto &v : vars)
        m = std::max(m, v.getFitterIndex());
    return m;
}

std::ostream &operator<<(std::ostream &o, const Variable &var) {
    o << var.getName() << ": " << var.getValue() << " +/- " << var.getError();
    if(!var.IsFixed())
        o << " [" << var.getLowerLimit() << ", " << var.getUpperLimit() << "]";
    if(var.getIndex() >= 0)
        o << " GooFit index: " << var.getIndex();
    if(var.getFitterIndex() >= 0)
        o << " Fitter index: " << var.getFitterIndex();
    return o;
}

std::ostream &operator<<(std::ostream &o, const Observable &var) {
    o << var.getName() << ": " << var.getValue() << " (" << var.getNumBins() << " bins)";
    o << " [" << var.getLowerLimit() << ", " << var.getUpperLimit() << "]";
    if(var.getIndex() >= 0)
        o << " GooFit index: " << var.getIndex();
    return o;
}

std::istream &operator>>(std::istream &i, Observable &var) {
    fptype v;
    i >> v;
    var.setValue(v);
    return i;
}

// ---------------------------------------------------------------- DataSet

DataSet::DataSet(const Observable &var, std::string n)
    : name(std::move(n))
    , observables({var}) {
    generateName();
}
DataSet::DataSet(const std::vector<Observable> &vars, std::string n)
    : name(std::move(n))
    , observables(vars) {
    generateName();
}
DataSet::DataSet(const std::set<Observable> &vars, std::string n)
    : name(std::move(n))
    , observables(vars.begin(), vars.end()) {
    generateName();
}
DataSet::DataSet(std::initializer_list<Observable> vars, std::string n)
    : name(std::move(n))
    , observables(vars) {
    generateName();
}

void DataSet::addWeightedEvent(fptype) { throw GeneralError("AddWeightedEvent not implemented for this type of DataSet"); }

std::vector<fptype> DataSet::getCurrentValues() const {
    std::vector<fptype> v;
    for(const auto &o : observables)
        v.push_back(o.getValue());
    return v;
}

size_t DataSet::indexOfVariable(const Observable &var) const {
    for(size_t i = 0; i < observables.size(); i++)
        if(var == observables[i])
            return i;
    throw GeneralError("Invalid variable access into dataset: {}", var.getName());
}

void DataSet::checkAllVars() const {
    for(const Observable &v : observables)
        if(!v)
            throw OutOfRange(v.getName(), v.getValue(), v.getLowerLimit(), v.getUpperLimit());
}

void DataSet::generateName() {
    if(!name.empty())
        return;
    name = "data";
    for(const auto &o : observables)
        name += "_" + o.getName();
}

// ---------------------------------------------------------------- UnbinnedDataSet

UnbinnedDataSet::UnbinnedDataSet(const Observable &var, std::string n)
    : DataSet(var, std::move(n)) {
    data.resize(1);
}
UnbinnedDataSet::UnbinnedDataSet(const std::vector<Observable> &vars, std::string n)
    : DataSet(vars, std::move(n)) {
    data.resize(vars.size());
}
UnbinnedDataSet::UnbinnedDataSet(const std::set<Observable> &vars, std::string n)
    : DataSet(vars, std::move(n)) {
    data.resize(vars.size());
}
UnbinnedDataSet::UnbinnedDataSet(std::initializer_list<Observable> vars, std::string n)
    : DataSet(vars, std::move(n)) {
    data.resize(vars.size());
}

void UnbinnedDataSet::addEvent() {
    checkAllVars();
    for(size_t i = 0; i < observables.size(); i++)
        data.at(i).push_back(observables[i].getValue());
    numEventsAdded++;
}

void UnbinnedDataSet::clear() {
    for(auto &d : data)
        d.clear();
    numEventsAdded = 0;
}

fptype UnbinnedDataSet::getValue(const Observable &var, size_t idx) const {
    if(idx >= getNumEvents())
        throw GeneralError("UnbinnedDataSet: Attempted to find {} in event {} when only {} events exits", var.getName(),
                           idx, getNumEvents());
    return data[indexOfVariable(var)].at(idx);
}

void UnbinnedDataSet::loadEvent(size_t idx) {
    for(size_t i = 0; i < observables.size(); i++)
        observables[i].setValue(data.at(i).at(idx));
}

void UnbinnedDataSet::setValueForAllEvents(const Observable &var) {
    size_t k = indexOfVariable(var);
    std::fill(data[k].begin(), data[k].end(), var.getValue());
}

const std::vector<fptype> &UnbinnedDataSet::getColumn(const Observable &var) const {
    return data.at(indexOfVariable(var));
}

void UnbinnedDataSet::appendColumns(const std::vector<const fptype *> &cols, size_t n) {
    if(cols.size() != observables.size())
        throw GeneralError("appendColumns: {} columns for {} observables", cols.size(), observables.size());
    for(size_t i = 0; i < cols.size(); i++) {
        const fptype lo = observables[i].getLowerLimit(), hi = observables[i].getUpperLimit();
        for(size_t e = 0; e < n; e++)
            if(lo != hi && (cols[i][e] < lo || cols[i][e] > hi))
                throw OutOfRange(observables[i].getName(), cols[i][e], lo, hi);
    }
    for(size_t i = 0; i < cols.size(); i++)
        data[i].insert(data[i].end(), cols[i], cols[i] + n);
    numEventsAdded += n;
}

// Grid of bin centres lo + (i + 0.5) * step, the first observable varying fastest
// (UnbinnedDataSet::fillWithGrid, src/goofit/UnbinnedDataSet.cpp:67-105).
void UnbinnedDataSet::fillWithGrid() {
    for(const auto &d : data)
        if(!d.empty())
            throw GeneralError("The dataset must be empty to fill with grid!");
    std::vector<size_t> grid_vars;
    int evt_num = -1;
    for(size_t i = 0; i < observables.size(); i++) {
        if(observables[i].isEventNumber())
            evt_num = (int)i;
        else
            grid_vars.push_back(i);
    }
    size_t total = 1;
    for(size_t g : grid_vars)
        total *= observables[g].getNumBins();
    std::vector<size_t> idx(grid_vars.size(), 0);
    const fptype first = evt_num >= 0 ? observables[evt_num].getValue() : 0;
    for(size_t n = 0; n < total; n++) {
        for(size_t k = 0; k < grid_vars.size(); k++) {
            Observable &v = observables[grid_vars[k]];
            double step   = (v.getUpperLimit() - v.getLowerLimit()) / v.getNumBins();
            v.setValue(v.getLowerLimit() + (idx[k] + 0.5) * step);
        }
        if(evt_num >= 0)
            observables[evt_num].setValue(first + (fptype)n);
        addEvent();
        for(size_t k = 0; k < grid_vars.size(); k++) {
            if(++idx[k] < observables[grid_vars[k]].getNumBins())
                break;
            idx[k] = 0;
        }
    }
}

}  // namespace GooFit
