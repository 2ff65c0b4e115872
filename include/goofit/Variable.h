// Observables and fit parameters of the GooFit API (include/goofit/Variable.h of
// the reference).  Copies share state: every field lives behind a shared_ptr, so
// a Variable passed by value into a PDF is the same parameter the fitter moves,
// and equality means "same underlying object".
#pragma once

#include <iosfwd>
#include <limits>
#include <memory>
#include <string>
#include <vector>

#include <goofit/GlobalCudaDefines.h>

namespace GooFit {

class Params;
class BatchFitter;

/// A named, limited, indexable quantity (min == max means "no limits").
class Indexable {
  public:
    Indexable(std::string n, fptype val = 0)
        : name_(std::move(n))
        , value_(std::make_shared<fptype>(val)) {}
    Indexable(std::string n, fptype val, fptype lo, fptype hi)
        : name_(std::move(n))
        , value_(std::make_shared<fptype>(val))
        , lower_(std::make_shared<fptype>(lo))
        , upper_(std::make_shared<fptype>(hi)) {}
    virtual ~Indexable() = default;

    bool isEqualNameValLimits(const Indexable &other) const {
        return name_ == other.name_ && getValue() == other.getValue() && getLowerLimit() == other.getLowerLimit()
               && getUpperLimit() == other.getUpperLimit();
    }

    int getIndex() const { return *index_; }
    void setIndex(int v) { *index_ = v; }
    const std::string &getName() const { return name_; }
    fptype getValue() const { return *value_; }
    void setValue(fptype v) { *value_ = v; }
    fptype getUpperLimit() const { return *upper_; }
    void setUpperLimit(fptype v) { *upper_ = v; }
    fptype getLowerLimit() const { return *lower_; }
    void setLowerLimit(fptype v) { *lower_ = v; }

    const Indexable &operator=(const fptype &v) {
        setValue(v);
        return *this;
    }
    operator fptype() const { return getValue(); }
    bool operator<(const Indexable &o) const { return value_.get() < o.value_.get(); }
    bool operator==(const Indexable &o) const { return value_.get() == o.value_.get(); }
    bool operator!=(const Indexable &o) const { return value_.get() != o.value_.get(); }
    /// In range (or unlimited)
    explicit operator bool() const {
        return getUpperLimit() == getLowerLimit() || (getValue() <= getUpperLimit() && getValue() >= getLowerLimit());
    }

  protected:
    std::string name_;
    std::shared_ptr<fptype> value_;
    std::shared_ptr<fptype> lower_{std::make_shared<fptype>(0)};
    std::shared_ptr<fptype> upper_{std::make_shared<fptype>(0)};
    std::shared_ptr<int> index_{std::make_shared<int>(-1)};
};

/// An observable: a column of a DataSet, with a binning for grids and plots.
class Observable : public Indexable {
  public:
    Observable(std::string name, fptype lo, fptype hi)
        : Indexable(std::move(name), 0, lo, hi) {}
    ~Observable() override = default;

    void setNumBins(size_t n) { *numbins_ = n; }
    size_t getNumBins() const { return *numbins_; }
    fptype getBinSize() const { return (getUpperLimit() - getLowerLimit()) / getNumBins(); }
    bool isEventNumber() const { return is_event_number_; }

    const Observable &operator=(const fptype &v) {
        setValue(v);
        return *this;
    }

  protected:
    std::shared_ptr<size_t> numbins_{std::make_shared<size_t>(100)};
    bool is_event_number_{false};
};

/// A fit parameter.
class Variable : public Indexable {
  public:
    /// Key to the blinding offset (only the fitter may read it).
    class Key {
        friend Params;
        friend BatchFitter;
        Key() = default;
    };

    Variable(const Variable &) = default;
    /// Constant.
    Variable(std::string n, fptype v)
        : Indexable(std::move(n), v, v + 0.01, v - 0.01)
        , error_(std::make_shared<fptype>(0.002)) {
        *fixed_ = true;
    }
    /// Value and error; fixed, as in the reference (include/goofit/Variable.h:170-175).
    Variable(std::string n, fptype v, fptype err)
        : Indexable(std::move(n), v, 0, 0)
        , error_(std::make_shared<fptype>(err)) {
        *fixed_ = true;
    }
    /// Free, with limits; error 10% of the range.
    Variable(std::string n, fptype v, fptype lo, fptype hi)
        : Indexable(std::move(n), v, lo, hi)
        , error_(std::make_shared<fptype>(0.1 * (hi - lo))) {}
    /// Value, error, limits and fixed flag (lo == hi: no limits).
    Variable(std::string n, fptype v, fptype err, fptype lo, fptype hi, bool fix = false)
        : Indexable(std::move(n), v, lo, hi)
        , error_(std::make_shared<fptype>(err))
        , fixed_(std::make_shared<bool>(fix)) {}
    ~Variable() override = default;

    bool isEqualNameValLimitsErrBlindFixed(const Variable &o) const {
        return isEqualNameValLimits(o) && getError() == o.getError() && *blind_ == *o.blind_ && IsFixed() == o.IsFixed();
    }

    fptype getError() const { return *error_; }
    void setError(fptype v) { *error_ = v; }
    int getFitterIndex() const { return *fitter_index_; }
    void setFitterIndex(int v) { *fitter_index_ = v; }
    bool getChanged() const { return *changed_; }
    void setChanged(bool v = true) { *changed_ = v; }
    bool IsFixed() const { return *fixed_; }
    void setFixed(bool f) { *fixed_ = f; }
    /// The fitted value is the stored value minus this offset; cannot be read back.
    void setBlind(fptype v) { *blind_ = v; }
    fptype getBlind(const Key &) { return *blind_; }

    const Variable &operator=(const fptype &v) {
        setValue(v);
        return *this;
    }

  protected:
    std::shared_ptr<fptype> error_;
    std::shared_ptr<fptype> blind_{std::make_shared<fptype>(0)};
    std::shared_ptr<bool> changed_{std::make_shared<bool>(true)};
    std::shared_ptr<bool> fixed_{std::make_shared<bool>(false)};
    std::shared_ptr<int> fitter_index_{std::make_shared<int>(-1)};
};

/// The row number of an event (used by Amp3Body to index its wave cache).
class EventNumber : public Observable {
  public:
    static constexpr fptype maxint{static_cast<fptype>(1L << std::numeric_limits<fptype>::digits)};
    EventNumber(std::string name, fptype lo = 0, fptype hi = EventNumber::maxint)
        : Observable(std::move(name), lo, hi) {
        is_event_number_ = true;
        *value_          = 0;
    }
    ~EventNumber() override = default;
    const EventNumber &operator=(const fptype &v) {
        setValue(v);
        return *this;
    }
};

int max_index(const std::vector<Variable> &vars);
int max_index(const std::vector<Observable> &vars);
int max_fitter_index(const std::vector<Variable> &vars);

std::ostream &operator<<(std::ostream &o, const Variable &var);
std::ostream &operator<<(std::ostream &o, const Observable &var);
std::istream &operator>>(std::istream &i, Observable &var);

}  // namespace GooFit
