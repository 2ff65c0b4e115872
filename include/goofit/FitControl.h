// Metric selection of the GooFit API (include/goofit/FitControl.h of the reference).
#pragma once

#include <cstddef>
#include <string>

namespace GooFit {

enum class EvalFunc : size_t { Eval = 0, NLL, Prob, BinAvg, BinWithError, Chisq };

inline const char *evalfunc_to_string(EvalFunc val) {
    static const char *names[]
        = {"ptr_to_Eval", "ptr_to_NLL", "ptr_to_Prob", "ptr_to_BinAvg", "ptr_to_BinWithError", "ptr_to_Chisq"};
    return names[static_cast<size_t>(val)];
}

class FitControl {
  public:
    FitControl(bool bin, EvalFunc mn)
        : binned(bin)
        , metricFunc(mn) {}
    virtual ~FitControl() = default;
    bool binnedFit() const { return binned; }
    bool binErrors() const { return errorsOnBins; }
    bool metricIsPdf() const { return !errorsOnBins; }
    EvalFunc getMetric() const { return metricFunc; }
    std::string getName() const { return evalfunc_to_string(getMetric()); }

  protected:
    bool errorsOnBins{false};

  private:
    bool binned;
    EvalFunc metricFunc;
};

class UnbinnedNllFit : public FitControl {
  public:
    UnbinnedNllFit()
        : FitControl(false, EvalFunc::NLL) {}
};
class BinnedNllFit : public FitControl {
  public:
    BinnedNllFit()
        : FitControl(true, EvalFunc::BinAvg) {}
};
class BinnedErrorFit : public FitControl {
  public:
    BinnedErrorFit()
        : FitControl(true, EvalFunc::BinWithError) {
        errorsOnBins = true;
    }
};
class BinnedChisqFit : public FitControl {
  public:
    BinnedChisqFit()
        : FitControl(true, EvalFunc::Chisq) {}
};
class ProbFit : public FitControl {
  public:
    ProbFit()
        : FitControl(false, EvalFunc::Prob) {}
};
class EvalFit : public FitControl {
  public:
    EvalFit()
        : FitControl(false, EvalFunc::Eval) {}
};

}  // namespace GooFit
