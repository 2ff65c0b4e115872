// Minuit parameters of a PDF (include/goofit/fitting/Params.h of the reference).
#pragma once

#include <vector>

#include <Minuit2/FCNBase.h>
#include <Minuit2/MnUserParameterState.h>
#include <Minuit2/MnUserParameters.h>
#include <goofit/Variable.h>

namespace Minuit2 = ROOT::Minuit2;

namespace GooFit {

class PdfBase;
class FCN;

class Params : public Minuit2::MnUserParameters {
    friend FCN;

  protected:
    std::vector<Variable> vars_;
    PdfBase *pdf_;
    size_t num_;
    bool do_record_{false};
    std::vector<std::vector<double>> recorded_;
    // from_minuit_vector's two tree walks (PdfBase::updateVariable per variable,
    // then PdfBase::updateParameters), done once per layout of the flattened
    // arrays: the parameters each variable sets, and the (parameter, slot) copies
    std::vector<std::vector<Variable>> targets_;
    std::vector<std::pair<Variable, unsigned int>> slots_;
    unsigned long long plan_epoch_{~0ull};
    size_t plan_nhost_{0};

  public:
    using MnUserParameters::MnUserParameters;
    Params(PdfBase &pdf);

    std::vector<Variable> GetGooFitParams() const { return vars_; }
    /// Copy values and errors of a Minuit state back into the Variables.
    void SetGooFitParams(const Minuit2::MnUserParameterState &input);
    size_t size() const { return vars_.size(); }
    std::vector<double> make_minuit_vector() const;
    /// Push a Minuit vector into the PDF (changed flags, blinding, device upload).
    void from_minuit_vector(const std::vector<double> &values, bool force_changed = false);
    void set_record(bool do_record = true) { do_record_ = do_record; }
    std::vector<std::vector<double>> get_recorded() const { return recorded_; }
};

}  // namespace GooFit
