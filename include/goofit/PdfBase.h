// PdfBase: the node of a PDF tree (include/goofit/PdfBase.h of the reference).
//
// A PDF registers its parameters, constants and observables and one function
// code (the reference registers a device function pointer instead).
// setIndices() flattens the whole tree into the four arrays of Engine.h in the
// reference's layout (src/PDFs/PdfBase.cu:268-311); the GPU engine then
// evaluates that tree.
#pragma once

#include <iosfwd>
#include <memory>
#include <string>
#include <vector>
#include <utility>

#include <goofit/Engine.h>
#include <goofit/Error.h>
#include <goofit/GlobalCudaDefines.h>
#include <goofit/Variable.h>

namespace ROOT {
namespace Minuit2 {
class FunctionMinimum;
}  // namespace Minuit2
}  // namespace ROOT

namespace GooFit {

class FitControl;
class DataSet;
class BinnedDataSet;
class UnbinnedDataSet;

namespace detail {
inline void filter_arguments(std::vector<Observable> &, std::vector<Variable> &) {}
template <typename... Args>
void filter_arguments(std::vector<Observable> &ob, std::vector<Variable> &va, const Observable &o, Args... args);
template <typename... Args>
void filter_arguments(std::vector<Observable> &ob, std::vector<Variable> &va, const EventNumber &o, Args... args);
template <typename... Args>
void filter_arguments(std::vector<Observable> &ob, std::vector<Variable> &va, const Variable &v, Args... args);
template <typename... Args>
void filter_arguments(std::vector<Observable> &ob, std::vector<Variable> &va, const Observable &o, Args... args) {
    ob.push_back(o);
    filter_arguments(ob, va, args...);
}
template <typename... Args>
void filter_arguments(std::vector<Observable> &ob, std::vector<Variable> &va, const EventNumber &o, Args... args) {
    ob.push_back(o);
    filter_arguments(ob, va, args...);
}
template <typename... Args>
void filter_arguments(std::vector<Observable> &ob, std::vector<Variable> &va, const Variable &v, Args... args) {
    va.push_back(v);
    filter_arguments(ob, va, args...);
}
}  // namespace detail

class PdfBase {
    friend std::ostream &operator<<(std::ostream &, const PdfBase &);

  protected:
    /// Mark the flattened tree for upload (the reference copies every array to the device here).
    void pre_run();
    /// Mark the parameters for upload (reference: host_parameters.smart_sync).
    void pre_call();
    /// Append this node's blocks, then its components' (depth first).
    virtual void populateArrays();
    /// [lower, upper, bins] per observable, for the grid integral.
    void generateNormRange();

    void registerParameter(Variable var);
    void unregisterParameter(Variable var);
    void registerConstant(fptype value);
    /// Register the function this PDF evaluates (a gf_func code of include/goofit_hip.h).
    void registerFunction(std::string name, int function_code);
    void registerObservable(Observable obs);
    /// Set the normalisation of this node and every component (and upload unless subpdf).
    void recursiveSetNormalization(fptype norm = 1.0, bool subpdf = false);

  public:
    template <typename... Args>
    explicit PdfBase(std::string pdf_name, std::string n, Args... args)
        : pdf_name_(std::move(pdf_name))
        , name(std::move(n)) {
        std::vector<Observable> obs;
        std::vector<Variable> vars;
        detail::filter_arguments(obs, vars, args...);
        for(auto &o : obs)
            registerObservable(o);
        for(auto &v : vars)
            registerParameter(v);
    }
    virtual ~PdfBase() = default;

    virtual double calculateNLL()  = 0;
    virtual fptype normalize()     = 0;

    void initializeIndices();
    void copyParams();

    std::string getName() const { return name; }
    std::vector<Observable> getObservables() const;
    std::vector<Variable> getParameters() const;
    Variable *getParameterByName(std::string n);

    /// Upload a DataSet as SoA event columns (and re-flatten the tree).
    void setData(DataSet *data);
    DataSet *getData() { return data_; }

    virtual void setFitControl(std::shared_ptr<FitControl>) = 0;
    virtual bool hasAnalyticIntegral() const { return false; }

    /// Fill the (1-D) dataset with `events` values sampled from the PDF on its grid.
    void fillMCDataSimple(size_t events, unsigned int seed = 0);

    ROOT::Minuit2::FunctionMinimum fitTo(DataSet *data, int verbosity = 3);
    ROOT::Minuit2::FunctionMinimum fit(int verbosity = 3);

    unsigned int getFunctionIndex() const { return functionIdx; }
    unsigned int getParameterIndex() const { return parametersIdx; }

    void setNormalization(const fptype &v);
    fptype getNormalization() const { return cachedNormalization; }

    void setIntegrationFineness(int i);
    bool parametersChanged() const;
    void updateVariable(const Variable &v, fptype newValue);
    void updateParameters();
    /// Every (Variable, host_parameters slot) pair that updateParameters() copies,
    /// this PDF's first and then its components' in the same recursion, so that a
    /// caller that sets parameters many times can copy them in a flat loop.
    void parameterSlots(std::vector<std::pair<Variable, unsigned int>> &out) const;
    /// Every parameter of the tree that updateVariable(var, ...) sets (the same
    /// variable or one of the same name), in the order it visits them.
    void matchingParameters(const Variable &var, std::vector<Variable> &out) const;

    void SigGenSetIndices() {
        setupObservables();
        setIndices();
    }
    /// Number this PDF's observables 0..k-1 (the event column of each).
    void setupObservables();
    virtual void recursiveSetIndices();
    virtual void setIndices();

    void setCommonNorm(bool v = true) { commonNorm = v; }
    void setSeparateNorm(bool v = true) { separateNorm = v; }
    bool getCommonNorm() const { return commonNorm; }
    bool getSeparateNorm() const { return separateNorm; }
    std::vector<PdfBase *> getComponents() { return components; }
    std::string getPdfName() const { return pdf_name_; }
    bool areParamsandConstantsEqualByVal(const PdfBase &other) const;

  protected:
    DataSet *data_ = nullptr;
    std::string reflex_name_;
    int function_code_{GF_FN_NONE};

    fptype numEvents{0};
    unsigned int numEntries{0};
    std::vector<fptype> normRanges;

    unsigned int parameters{0};
    fptype cachedNormalization{1.0};

    std::shared_ptr<FitControl> fitControl;
    std::vector<Observable> observablesList;
    std::vector<Variable> parametersList;
    std::vector<fptype> constantsList;
    std::vector<PdfBase *> components;

    int integrationBins{-1};
    bool properlyInitialised{true};

    unsigned int functionIdx{0};
    unsigned int parametersIdx{0};
    unsigned int constantsIdx{0};
    unsigned int observablesIdx{0};
    unsigned int normalIdx{0};

    bool commonNorm{false};
    bool separateNorm{false};

  private:
    std::string pdf_name_;
    std::string name;
};

std::ostream &operator<<(std::ostream &, const PdfBase &);

}  // namespace GooFit
