// MIGRAD driver of a PDF (include/goofit/fitting/FitManagerMinuit2.h of the reference).
#pragma once

#include <utility>
#include <vector>

#include <Minuit2/FunctionMinimum.h>
#include <Minuit2/MnScan.h>
#include <goofit/fitting/FCN.h>

namespace GooFit {

class PdfBase;

enum class FitErrors : int { Valid = 0, NotRun = 50, InValid };

class FitManagerMinuit2 {
  public:
    FitManagerMinuit2(PdfBase *dat);
    ROOT::Minuit2::FunctionMinimum fit();
    void setMaxCalls(unsigned int max_calls = 0) { maxfcn_ = max_calls; }
    Params *getParams() { return &upar_; }
    FCN *getFCN() { return &fcn_; }
    ROOT::Minuit2::MnScan getMnScan();
    operator bool() const { return retval_ == FitErrors::Valid; }
    operator int() const { return static_cast<int>(retval_); }
    void setVerbosity(int value) { verbosity = value; }
    int getVerbosity() const { return verbosity; }
    std::vector<std::pair<double, double>> getMinosErrors() const { return minos_errors; }
    void setMinos(bool minos_flag = true) { minos = minos_flag; }
    void setStrategy(int strat) { m_strategy = strat; }
    /// MIGRAD tolerance (EDM target 0.002 * tolerance * Up; Minuit2 default 0.1).
    void setTolerance(double tol) { tolerance_ = tol; }
    /// Wall time of the last fit() in seconds.
    double getLastFitTime() const { return last_time_; }
    /// Hand the points of each derivative step to the GPU together (default on):
    /// the same MIGRAD with the same function values, fewer launches.
    void setBatchedDerivatives(bool on) { fcn_.setBatched(on); }
    bool getBatchedDerivatives() const { return fcn_.isBatched(); }

  private:
    Params upar_;
    FCN fcn_;
    unsigned int maxfcn_{0};
    FitErrors retval_{FitErrors::NotRun};
    int verbosity{3};
    std::vector<std::pair<double, double>> minos_errors;
    bool minos{false};
    int m_strategy{1};
    double tolerance_{0.1};
    double last_time_{0};
};

}  // namespace GooFit
