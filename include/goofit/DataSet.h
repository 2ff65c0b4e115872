// DataSet base class of the GooFit API (include/goofit/DataSet.h of the reference).
#pragma once

#include <initializer_list>
#include <set>
#include <string>
#include <vector>

#include <goofit/Error.h>
#include <goofit/GlobalCudaDefines.h>
#include <goofit/Variable.h>

namespace GooFit {

class DataSet {
  public:
    DataSet(const Observable &var, std::string n = "");
    DataSet(const std::vector<Observable> &vars, std::string n = "");
    DataSet(const std::set<Observable> &vars, std::string n = "");
    DataSet(std::initializer_list<Observable> vars, std::string n = "");
    virtual ~DataSet() = default;

    /// Append the current values of the observables (must increment numEventsAdded).
    virtual void addEvent() = 0;
    virtual void addWeightedEvent(fptype weight);

    /// Set the observables to the given values, then addEvent().
    template <typename... Args>
    void addEvent(fptype value, Args... args) {
        std::vector<fptype> values{value, static_cast<fptype>(args)...};
        if(values.size() != observables.size())
            throw GeneralError("You must pass the correct number of values ({}) to addEvent", observables.size());
        for(size_t i = 0; i < values.size(); i++)
            observables[i].setValue(values[i]);
        addEvent();
    }

    const std::vector<Observable> &getObservables() const { return observables; }
    size_t numVariables() const { return observables.size(); }
    size_t getNumEvents() const { return numEventsAdded; }
    std::string getName() const { return name; }

  protected:
    std::vector<fptype> getCurrentValues() const;
    size_t indexOfVariable(const Observable &var) const;
    /// Throw if an observable is outside its limits.
    void checkAllVars() const;
    size_t numEventsAdded{0};

  private:
    void generateName();
    std::string name;

  protected:
    std::vector<Observable> observables;
};

}  // namespace GooFit
