// BinnedDataSet (src/goofit/BinnedDataSet.cpp of the reference).  Bin numbers
// must match the reference bit for bit: the same operation order below.
#include <goofit/BinnedDataSet.h>
#include <goofit/Error.h>

#include <cmath>
#include <functional>
#include <iostream>
#include <numeric>

namespace GooFit {

BinnedDataSet::BinnedDataSet(const Observable &var, std::string n)
    : DataSet(var, n) {
    collectBins();
    binvalues.resize(getNumBins());
}

BinnedDataSet::BinnedDataSet(const std::vector<Observable> &vars, std::string n)
    : DataSet(vars, n) {
    collectBins();
    binvalues.resize(getNumBins());
}

BinnedDataSet::BinnedDataSet(const std::set<Observable> &vars, std::string n)
    : DataSet(vars, n) {
    collectBins();
    binvalues.resize(getNumBins());
}

BinnedDataSet::BinnedDataSet(std::initializer_list<Observable> vars, std::string n)
    : DataSet(vars, n) {
    collectBins();
    binvalues.resize(getNumBins());
}

// BinnedDataSet::addEvent, BinnedDataSet.cpp:37-42
void BinnedDataSet::addEvent() {
    checkAllVars();
    const size_t ibin = getBinNumber();
    binvalues.at(ibin) += 1;
    numEventsAdded++;
}

void BinnedDataSet::addWeightedEvent(double weight) {
    checkAllVars();
    const size_t ibin = getBinNumber();
    binvalues.at(ibin) += weight;
    numEventsAdded++;
}

void BinnedDataSet::collectBins() {
    binsizes.clear();
    for(const Observable &var : observables)
        binsizes.push_back(var.getNumBins());
}

size_t BinnedDataSet::getBinNumber() const { return localToGlobal(convertValuesToBins(getCurrentValues())); }

// BinnedDataSet.cpp:63-74: global = sum_i local_i * prod_{j<i} n_j (unsigned arithmetic)
size_t BinnedDataSet::localToGlobal(const std::vector<size_t> &locals) const {
    unsigned int prior = 1;
    unsigned int ret   = 0;
    for(size_t i = 0; i < observables.size(); i++) {
        ret += locals[i] * prior;
        prior *= binsizes[i];
    }
    return ret;
}

std::vector<size_t> BinnedDataSet::globalToLocal(size_t global) const {
    std::vector<size_t> locals;
    for(size_t i = 0; i < observables.size(); i++) {
        locals.push_back(global % binsizes[i]);
        global /= binsizes[i];
    }
    return locals;
}

// BinnedDataSet.cpp:92-100: size * (local + 0.5) + lower
fptype BinnedDataSet::getBinCenter(size_t ivar, size_t bin) const {
    const size_t local = globalToLocal(bin).at(ivar);
    fptype ret         = getBinSize(ivar);
    ret *= (local + 0.5);
    ret += observables[ivar].getLowerLimit();
    return ret;
}

fptype BinnedDataSet::getBinCenter(const Observable &var, size_t bin) const {
    return getBinCenter(indexOfVariable(var), bin);
}

fptype BinnedDataSet::getBinSize(size_t ivar) const {
    return (observables.at(ivar).getUpperLimit() - observables[ivar].getLowerLimit()) / binsizes[ivar];
}

fptype BinnedDataSet::getBinVolume(size_t) const {
    fptype ret = 1;
    for(size_t i = 0; i < observables.size(); i++)
        ret *= getBinSize(i);
    return ret;
}

fptype BinnedDataSet::getBinError(size_t bin) const {
    if(binerrors.empty())
        return std::sqrt(binvalues.at(bin));
    return binerrors.at(bin);
}

void BinnedDataSet::setBinError(unsigned int bin, fptype error) {
    if(binerrors.empty())
        binerrors.resize(binvalues.size());
    binerrors.at(bin) = error;
}

size_t BinnedDataSet::getNumBins() const {
    return std::accumulate(binsizes.begin(), binsizes.end(), (size_t)1, std::multiplies<size_t>());
}

// BinnedDataSet.cpp:135-137 accumulates into an int (0 literal): kept
fptype BinnedDataSet::getNumWeightedEvents() const { return std::accumulate(binvalues.begin(), binvalues.end(), 0); }

// BinnedDataSet.cpp:139-160: clamp into [lower, upper] (a value equal to the upper
// limit lands in bin n, outside the histogram), floor((v - lower) / size)
std::vector<size_t> BinnedDataSet::convertValuesToBins(const std::vector<fptype> &vals) const {
    if(vals.size() != observables.size())
        throw GeneralError("Incorrect number of bins {} for {} variables", vals.size(), observables.size());
    std::vector<size_t> locals;
    for(size_t i = 0; i < observables.size(); i++) {
        const fptype cur = vals[i];
        const fptype bet = std::min(std::max(cur, observables[i].getLowerLimit()), observables[i].getUpperLimit());
        if(cur != bet)
            std::cerr << "Warning: Value " << cur << " outside " << observables[i].getName() << " range ["
                      << observables[i].getLowerLimit() << "," << observables[i].getUpperLimit() << "] - clamping to "
                      << bet << std::endl;
        locals.push_back(static_cast<size_t>(std::floor((bet - observables[i].getLowerLimit()) / getBinSize(i))));
    }
    return locals;
}

}  // namespace GooFit
