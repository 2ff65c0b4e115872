// BinnedDataSet: a histogram over the observables' bins, stored as one flat
// content array indexed by the global bin number (first observable fastest).
//
// The axis table (range, bin count, stride of each observable) is built once at
// construction; every conversion between values, local bins and global bins is
// per-axis integer arithmetic on that table.  Three expressions are fixed by the
// reference because fits compare bin numbers and centres bit for bit:
//   local bin  = floor((clamp(v, lower, upper) - lower) / ((upper - lower) / n))
//                (src/goofit/BinnedDataSet.cpp:139-160; v == upper gives n, hazard A14),
//   global bin = sum of local * stride in 32-bit unsigned arithmetic (:63-74),
//   centre     = ((upper - lower) / n) * (local + 0.5) + lower (:92-100).
#include <goofit/BinnedDataSet.h>
#include <goofit/Error.h>

#include <cmath>
#include <cstdint>
#include <iostream>

namespace GooFit {

BinnedDataSet::BinnedDataSet(const Observable &var, std::string n)
    : DataSet(var, n) {
    buildAxes();
}

BinnedDataSet::BinnedDataSet(const std::vector<Observable> &vars, std::string n)
    : DataSet(vars, n) {
    buildAxes();
}

BinnedDataSet::BinnedDataSet(const std::set<Observable> &vars, std::string n)
    : DataSet(vars, n) {
    buildAxes();
}

BinnedDataSet::BinnedDataSet(std::initializer_list<Observable> vars, std::string n)
    : DataSet(vars, n) {
    buildAxes();
}

void BinnedDataSet::buildAxes() {
    axes.clear();
    size_t stride = 1;
    for(const Observable &o : observables) {
        axes.push_back(Axis{o.getLowerLimit(), o.getUpperLimit(), o.getNumBins(), stride});
        stride *= o.getNumBins();
    }
    binvalues.assign(stride, 0.0);
    binerrors.clear();
}

size_t BinnedDataSet::localBin(size_t a, fptype v) const {
    const Axis &ax  = axes[a];
    const fptype in = std::min(std::max(v, ax.lower), ax.upper);
    if(in != v)
        std::cerr << "Warning: Value " << v << " outside " << observables[a].getName() << " range [" << ax.lower << ","
                  << ax.upper << "] - clamping to " << in << std::endl;
    return static_cast<size_t>(std::floor((in - ax.lower) / getBinSize(a)));
}

size_t BinnedDataSet::getBinNumber() const {
    const std::vector<fptype> vals = getCurrentValues();
    if(vals.size() != axes.size())
        throw GeneralError("Incorrect number of bins {} for {} variables", vals.size(), axes.size());
    // 32-bit unsigned accumulation: a histogram of more than 2^32 bins wraps as in the reference
    uint32_t global = 0;
    for(size_t a = 0; a < axes.size(); a++)
        global += static_cast<uint32_t>(localBin(a, vals[a]) * static_cast<uint32_t>(axes[a].stride));
    return global;
}

void BinnedDataSet::addEvent() { addWeightedEvent(1.0); }

void BinnedDataSet::addWeightedEvent(double weight) {
    checkAllVars();
    binvalues.at(getBinNumber()) += weight;  // out of range (A14 in 1-D) throws std::out_of_range
    numEventsAdded++;
}

fptype BinnedDataSet::getBinSize(size_t ivar) const {
    const Axis &ax = axes.at(ivar);
    return (ax.upper - ax.lower) / ax.bins;
}

fptype BinnedDataSet::getBinCenter(size_t ivar, size_t bin) const {
    fptype centre = getBinSize(ivar);
    centre *= (localOf(ivar, bin) + 0.5);
    centre += axes[ivar].lower;
    return centre;
}

fptype BinnedDataSet::getBinCenter(const Observable &var, size_t bin) const {
    return getBinCenter(indexOfVariable(var), bin);
}

fptype BinnedDataSet::getBinVolume(size_t) const {
    fptype volume = 1;
    for(size_t a = 0; a < axes.size(); a++)
        volume *= getBinSize(a);
    return volume;
}

fptype BinnedDataSet::getBinError(size_t bin) const {
    return binerrors.empty() ? std::sqrt(binvalues.at(bin)) : binerrors.at(bin);
}

void BinnedDataSet::setBinError(unsigned int bin, fptype error) {
    if(binerrors.empty())
        binerrors.assign(binvalues.size(), 0.0);
    binerrors.at(bin) = error;
}

size_t BinnedDataSet::getNumBins() const { return binvalues.size(); }

std::vector<size_t> BinnedDataSet::getDiminsions() const {
    std::vector<size_t> out;
    for(const Axis &ax : axes)
        out.push_back(ax.bins);
    return out;
}

// The reference sums the contents with an int accumulator (std::accumulate from the
// literal 0, src/goofit/BinnedDataSet.cpp:135-137), truncating after every addition.
fptype BinnedDataSet::getNumWeightedEvents() const {
    int total = 0;
    for(fptype v : binvalues)
        total = static_cast<int>(total + v);
    return total;
}

}  // namespace GooFit
