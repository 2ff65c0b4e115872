// Column-stored event sample (include/goofit/UnbinnedDataSet.h of the reference).
#pragma once

#include <goofit/DataSet.h>

namespace GooFit {

class UnbinnedDataSet : public DataSet {
  public:
    using DataSet::addEvent;

    UnbinnedDataSet(const Observable &var, std::string n = "");
    UnbinnedDataSet(const std::vector<Observable> &vars, std::string n = "");
    UnbinnedDataSet(const std::set<Observable> &vars, std::string n = "");
    UnbinnedDataSet(std::initializer_list<Observable> vars, std::string n = "");
    ~UnbinnedDataSet() override = default;

    void addEvent() override;

    /// Delete all events.
    void clear();

    /// Replace the (empty) dataset by the bin-centre grid of its observables,
    /// first observable fastest; an EventNumber observable counts the points.
    void fillWithGrid();

    fptype getValue(const Observable &var, size_t idx) const;
    /// Set every observable to the values of event idx.
    void loadEvent(size_t idx);
    /// Set the column of var to var's current value for every event.
    void setValueForAllEvents(const Observable &var);

    /// Whole column of an observable (for bulk upload).
    const std::vector<fptype> &getColumn(const Observable &var) const;
    /// Append n events from columns ordered as getObservables() (bulk fill).
    void appendColumns(const std::vector<const fptype *> &cols, size_t n);

  private:
    std::vector<std::vector<fptype>> data;
};

}  // namespace GooFit
