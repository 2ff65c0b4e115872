// Binned data set of the GooFit API (include/goofit/BinnedDataSet.h of the
// reference): a histogram over the observables' bins, global bin number with
// the first observable fastest.
#pragma once

#include <initializer_list>
#include <set>
#include <string>
#include <vector>

#include <goofit/DataSet.h>

namespace GooFit {

class BinnedDataSet : public DataSet {
  public:
    using DataSet::addEvent;

    BinnedDataSet(const Observable &var, std::string n = "");
    BinnedDataSet(const std::vector<Observable> &vars, std::string n = "");
    BinnedDataSet(const std::set<Observable> &vars, std::string n = "");
    BinnedDataSet(std::initializer_list<Observable> vars, std::string n = "");
    ~BinnedDataSet() override = default;

    /// One entry in the bin of the observables' current values.
    void addEvent() override;
    void addWeightedEvent(double weight) override;

    fptype getBinContent(size_t bin) const { return binvalues.at(bin); }
    fptype getBinCenter(size_t ivar, size_t bin) const;
    fptype getBinCenter(const Observable &var, size_t bin) const;
    fptype getBinSize(size_t ivar) const;
    /// Bin of the observables' current values.
    size_t getBinNumber() const;
    fptype getBinVolume(size_t bin) const;
    /// sqrt(content) unless errors were set.
    fptype getBinError(size_t bin) const;
    std::vector<size_t> getDiminsions() const;
    size_t getNumBins() const;
    fptype getNumWeightedEvents() const;

    void setBinContent(unsigned int bin, fptype value) { binvalues.at(bin) = value; }
    void setBinError(unsigned int bin, fptype value);

  private:
    /// One histogram axis: the observable's range, its bin count, and the step of
    /// its local bin in the global bin number (the product of the counts of the
    /// axes before it: the first observable varies fastest).
    struct Axis {
        fptype lower, upper;
        size_t bins, stride;
    };
    /// Local bin of value v on axis a: v clamped into [lower, upper] (with a
    /// warning), then floor((v - lower) / bin size); v == upper gives `bins`.
    size_t localBin(size_t a, fptype v) const;
    /// Local bin of axis a inside global bin `bin`.
    size_t localOf(size_t a, size_t bin) const { return (bin / axes[a].stride) % axes[a].bins; }
    void buildAxes();

    std::vector<Axis> axes;
    std::vector<fptype> binvalues;
    std::vector<fptype> binerrors;  // empty until the first setBinError
};

}  // namespace GooFit
