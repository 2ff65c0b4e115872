// Python binding of the GooFit C++ API (the reference ships the same classes as the
// pybind11 module `goofit`, python/goofit.cpp and python/PDFs/**).  Everything here
// forwards to libgoofit.so; there is no numerics in this file.
#include <pybind11/complex.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>

#include <chrono>
#include <pybind11/stl.h>

#include <sstream>

#include <Minuit2/Minuit2.h>
#include <goofit/Engine.h>
#include <goofit/FitManager.h>
#include <goofit/PDFs/GooPdf.h>
#include <goofit/PDFs/basic/ExpGausPdf.h>
#include <goofit/PDFs/basic/ExpPdf.h>
#include <goofit/PDFs/basic/GaussianPdf.h>
#include <goofit/PDFs/basic/PolynomialPdf.h>
#include <goofit/PDFs/combine/AddPdf.h>
#include <goofit/PDFs/combine/ProdPdf.h>
#include <goofit/PDFs/physics/Amp3Body.h>
#include <goofit/PDFs/physics/Amp3Body_TD.h>
#include <goofit/PDFs/physics/ThreeGaussResolution.h>
#include <goofit/PDFs/physics/TruthResolution.h>
#include <goofit/PDFs/physics/DalitzProposal.h>
#include <goofit/fitting/BatchFitter.h>
#include <goofit/PDFs/physics/ResonancePdf.h>
#include <goofit/BinnedDataSet.h>
#include <goofit/FitControl.h>
#include <goofit/UnbinnedDataSet.h>
#include <goofit/Variable.h>

namespace py = pybind11;
using namespace GooFit;
namespace M2 = ROOT::Minuit2;

namespace {

template <class T>
std::string repr_of(const T &v) {
    std::ostringstream os;
    os << v;
    return os.str();
}

py::array_t<double> to_numpy(const std::vector<double> &v) {
    py::array_t<double> a(v.size());
    std::copy(v.begin(), v.end(), a.mutable_data());
    return a;
}

}  // namespace

PYBIND11_MODULE(_goofit, m) {
    m.doc() = "GooFit API on the MI355X likelihood engine";

    py::register_exception<GeneralError>(m, "GeneralError", PyExc_RuntimeError);

    // ------------------------------------------------------------ variables
    py::class_<Indexable>(m, "Indexable")
        .def_property_readonly("name", &Indexable::getName)
        .def_property("value", &Indexable::getValue, &Indexable::setValue)
        .def_property("lowerlimit", &Indexable::getLowerLimit, &Indexable::setLowerLimit)
        .def_property("upperlimit", &Indexable::getUpperLimit, &Indexable::setUpperLimit)
        .def_property("index", &Indexable::getIndex, &Indexable::setIndex)
        .def("getValue", &Indexable::getValue)
        .def("setValue", &Indexable::setValue)
        .def("getName", &Indexable::getName)
        .def("__bool__", [](const Indexable &v) { return (bool)v; })
        .def("__float__", [](const Indexable &v) { return v.getValue(); })
        .def("__eq__", [](const Indexable &a, const Indexable &b) { return a == b; })
        .def("__hash__", [](const Indexable &a) { return std::hash<std::string>()(a.getName()); });

    py::class_<Observable, Indexable>(m, "Observable")
        .def(py::init<std::string, fptype, fptype>(), py::arg("name"), py::arg("min"), py::arg("max"))
        .def_property("numbins", &Observable::getNumBins, &Observable::setNumBins)
        .def("getNumBins", &Observable::getNumBins)
        .def("setNumBins", &Observable::setNumBins)
        .def("getBinSize", &Observable::getBinSize)
        .def_property_readonly("bin_size", &Observable::getBinSize)
        .def_property_readonly("bin_centers",
                               [](const Observable &o) {
                                   std::vector<double> c(o.getNumBins());
                                   for(size_t i = 0; i < c.size(); i++)
                                       c[i] = o.getLowerLimit() + (i + 0.5) * o.getBinSize();
                                   return c;
                               })
        .def("isEventNumber", &Observable::isEventNumber)
        .def("__repr__", [](const Observable &o) { return repr_of(o); });

    py::class_<EventNumber, Observable>(m, "EventNumber")
        .def(py::init<std::string>(), py::arg("name"))
        .def(py::init<std::string, fptype, fptype>(), py::arg("name"), py::arg("min"), py::arg("max"));

    py::class_<Variable, Indexable>(m, "Variable")
        .def(py::init<std::string, fptype>(), py::arg("name"), py::arg("value"))
        .def(py::init<std::string, fptype, fptype>(), py::arg("name"), py::arg("value"), py::arg("error"))
        .def(py::init<std::string, fptype, fptype, fptype>(), py::arg("name"), py::arg("value"), py::arg("min"),
             py::arg("max"))
        .def(py::init<std::string, fptype, fptype, fptype, fptype, bool>(), py::arg("name"), py::arg("value"),
             py::arg("error"), py::arg("min"), py::arg("max"), py::arg("fixed") = false)
        .def_property("error", &Variable::getError, &Variable::setError)
        .def_property("fixed", &Variable::IsFixed, &Variable::setFixed)
        .def_property("fitterindex", &Variable::getFitterIndex, &Variable::setFitterIndex)
        .def_property_readonly("fitterIndex", &Variable::getFitterIndex)
        .def_property("blind", nullptr, &Variable::setBlind)
        .def("getError", &Variable::getError)
        .def("setBlind", &Variable::setBlind)
        .def("__repr__", [](const Variable &v) { return repr_of(v); });

    // ------------------------------------------------------------ data
    py::class_<DataSet>(m, "DataSet")
        .def("getNumEvents", &DataSet::getNumEvents)
        .def("getObservables", &DataSet::getObservables)
        .def("addEvent", [](DataSet &d) { d.addEvent(); })
        .def("getName", &DataSet::getName)
        .def_property_readonly("name", &DataSet::getName)
        .def_property_readonly("variables", &DataSet::getObservables)
        .def("__len__", &DataSet::getNumEvents);

    py::class_<UnbinnedDataSet, DataSet>(m, "UnbinnedDataSet")
        .def(py::init<const Observable &, std::string>(), py::arg("var"), py::arg("name") = "")
        .def(py::init<const std::vector<Observable> &, std::string>(), py::arg("vars"), py::arg("name") = "")
        .def(py::init([](py::args args) {
            std::vector<Observable> obs;
            for(auto a : args)
                obs.push_back(a.cast<Observable>());
            return new UnbinnedDataSet(obs);
        }))
        .def("getValue", &UnbinnedDataSet::getValue)
        .def("loadEvent", &UnbinnedDataSet::loadEvent)
        .def("fillWithGrid", &UnbinnedDataSet::fillWithGrid)
        .def("clear", &UnbinnedDataSet::clear)
        .def("setValueForAllEvents", &UnbinnedDataSet::setValueForAllEvents)
        .def("column", [](const UnbinnedDataSet &d, const Observable &o) { return to_numpy(d.getColumn(o)); })
        .def("from_columns",
             [](UnbinnedDataSet &d, std::vector<py::array_t<double, py::array::c_style | py::array::forcecast>> cols) {
                 std::vector<const double *> ptrs;
                 size_t n = cols.empty() ? 0 : (size_t)cols[0].size();
                 for(auto &c : cols) {
                     if((size_t)c.size() != n)
                         throw GeneralError("from_columns: columns of different lengths");
                     ptrs.push_back(c.data());
                 }
                 d.appendColumns(ptrs, n);
             })
        .def("__len__", &UnbinnedDataSet::getNumEvents)
        // python/goofit/UnbinnedDataSet.cpp of the reference: observables x events
        .def("to_numpy",
             [](const UnbinnedDataSet &d) {
                 const auto &obs = d.getObservables();
                 py::array_t<double> a({(py::ssize_t)obs.size(), (py::ssize_t)d.getNumEvents()});
                 for(size_t i = 0; i < obs.size(); i++) {
                     const auto &col = d.getColumn(obs[i]);
                     std::copy(col.begin(), col.end(), a.mutable_data() + i * d.getNumEvents());
                 }
                 return a;
             })
        .def("to_matrix",
             [](const UnbinnedDataSet &d) {
                 const auto &obs = d.getObservables();
                 py::array_t<double> a({(py::ssize_t)obs.size(), (py::ssize_t)d.getNumEvents()});
                 for(size_t i = 0; i < obs.size(); i++) {
                     const auto &col = d.getColumn(obs[i]);
                     std::copy(col.begin(), col.end(), a.mutable_data() + i * d.getNumEvents());
                 }
                 return a;
             })
        .def(
            "from_matrix",
            [](UnbinnedDataSet &d, py::array_t<double, py::array::c_style | py::array::forcecast> m, bool filter) {
                // rows = observables; one row fewer: the last observable counts the events
                auto obs          = d.getObservables();
                const size_t rows = (size_t)m.shape(0), n = (size_t)m.shape(1);
                if(rows != obs.size() && rows + 1 != obs.size())
                    throw GeneralError("from_matrix: {} rows for {} observables", rows, obs.size());
                const size_t first = d.getNumEvents();
                for(size_t j = 0; j < n; j++) {
                    for(size_t i = 0; i < rows; i++)
                        obs[i].setValue(m.at(i, j));
                    if(rows + 1 == obs.size())
                        obs.back().setValue((double)(first + j));
                    bool ok = true;
                    if(filter)
                        for(const auto &o : obs)
                            ok = ok && (bool)o;
                    if(ok)
                        d.addEvent();
                }
            },
            py::arg("matrix"), py::arg("filter") = false)
        .def(
            "from_numpy",
            [](UnbinnedDataSet &d, py::array_t<double, py::array::c_style | py::array::forcecast> m, bool filter) {
                auto obs = d.getObservables();
                for(py::ssize_t j = 0; j < m.shape(1); j++) {
                    for(py::ssize_t i = 0; i < m.shape(0); i++)
                        obs.at(i).setValue(m.at(i, j));
                    bool ok = true;
                    if(filter)
                        for(const auto &o : obs)
                            ok = ok && (bool)o;
                    if(ok)
                        d.addEvent();
                }
            },
            py::arg("array"), py::arg("filter") = false)
        .def("__getitem__", [](const UnbinnedDataSet &d, py::object idx) {
            const auto &obs = d.getObservables();
            py::array_t<double> a({(py::ssize_t)obs.size(), (py::ssize_t)d.getNumEvents()});
            for(size_t i = 0; i < obs.size(); i++) {
                const auto &col = d.getColumn(obs[i]);
                std::copy(col.begin(), col.end(), a.mutable_data() + i * d.getNumEvents());
            }
            return a.attr("__getitem__")(idx);
        });

    // python/goofit/BinnedDataSet.cpp of the reference
    py::class_<BinnedDataSet, DataSet>(m, "BinnedDataSet")
        .def(py::init<const Observable &, std::string>(), py::arg("var"), py::arg("name") = "")
        .def(py::init<const std::vector<Observable> &, std::string>(), py::arg("vars"), py::arg("name") = "")
        .def(py::init([](py::args args) {
            std::vector<Observable> obs;
            for(auto a : args)
                obs.push_back(a.cast<Observable>());
            return new BinnedDataSet(obs);
        }))
        .def("addEvent", [](BinnedDataSet &d) { d.addEvent(); })
        .def("addWeightedEvent", &BinnedDataSet::addWeightedEvent, py::arg("weight"))
        .def("getBinContent", &BinnedDataSet::getBinContent, py::arg("bin"))
        .def("setBinContent", &BinnedDataSet::setBinContent, py::arg("bin"), py::arg("value"))
        .def("getBinError", &BinnedDataSet::getBinError, py::arg("bin"))
        .def("setBinError", &BinnedDataSet::setBinError, py::arg("bin"), py::arg("value"))
        .def("getBinCenter", (fptype(BinnedDataSet::*)(size_t, size_t) const) & BinnedDataSet::getBinCenter)
        .def("getBinCenter", (fptype(BinnedDataSet::*)(const Observable &, size_t) const) & BinnedDataSet::getBinCenter)
        .def("getBinSize", &BinnedDataSet::getBinSize, py::arg("ivar"))
        .def("getBinNumber", &BinnedDataSet::getBinNumber)
        .def("getBinVolume", &BinnedDataSet::getBinVolume, py::arg("bin"))
        .def("getNumBins", &BinnedDataSet::getNumBins)
        .def("getNumWeightedEvents", &BinnedDataSet::getNumWeightedEvents)
        .def("getDiminsions", &BinnedDataSet::getDiminsions)
        .def_property_readonly("num_bins", &BinnedDataSet::getNumBins)
        .def_property_readonly("num_events", &BinnedDataSet::getNumEvents)
        .def_property_readonly("diminsions", &BinnedDataSet::getDiminsions)
        .def("__len__", &BinnedDataSet::getNumBins);

    // ------------------------------------------------------------ PDFs
    py::enum_<EvalFunc>(m, "EvalFunc")
        .value("Eval", EvalFunc::Eval)
        .value("NLL", EvalFunc::NLL)
        .value("Prob", EvalFunc::Prob)
        .value("BinAvg", EvalFunc::BinAvg)
        .value("BinWithError", EvalFunc::BinWithError)
        .value("Chisq", EvalFunc::Chisq);

    // python/goofit/FitControl.cpp of the reference
    py::class_<FitControl, std::shared_ptr<FitControl>>(m, "FitControl")
        .def("binnedFit", &FitControl::binnedFit)
        .def("binErrors", &FitControl::binErrors)
        .def("metricIsPdf", &FitControl::metricIsPdf)
        .def("getMetric", &FitControl::getMetric)
        .def("getName", &FitControl::getName);
    py::class_<UnbinnedNllFit, FitControl, std::shared_ptr<UnbinnedNllFit>>(m, "UnbinnedNllFit").def(py::init<>());
    py::class_<BinnedNllFit, FitControl, std::shared_ptr<BinnedNllFit>>(m, "BinnedNllFit").def(py::init<>());
    py::class_<BinnedErrorFit, FitControl, std::shared_ptr<BinnedErrorFit>>(m, "BinnedErrorFit").def(py::init<>());
    py::class_<BinnedChisqFit, FitControl, std::shared_ptr<BinnedChisqFit>>(m, "BinnedChisqFit").def(py::init<>());
    py::class_<ProbFit, FitControl, std::shared_ptr<ProbFit>>(m, "ProbFit").def(py::init<>());
    py::class_<EvalFit, FitControl, std::shared_ptr<EvalFit>>(m, "EvalFit").def(py::init<>());

    py::class_<PdfBase>(m, "PdfBase")
        .def("getName", &PdfBase::getName)
        .def("getParameters", &PdfBase::getParameters)
        .def("getObservables", &PdfBase::getObservables)
        .def("setData", &PdfBase::setData, py::keep_alive<1, 2>())
        .def("getData", &PdfBase::getData, py::return_value_policy::reference)
        .def("normalize", &PdfBase::normalize)
        .def("calculateNLL", &PdfBase::calculateNLL)
        .def("copyParams", &PdfBase::copyParams)
        .def("setIntegrationFineness", &PdfBase::setIntegrationFineness)
        .def("getFunctionIndex", &PdfBase::getFunctionIndex)
        .def("getParameterIndex", &PdfBase::getParameterIndex)
        .def("getNormalization", &PdfBase::getNormalization)
        .def("setupObservables", &PdfBase::setupObservables)
        .def("setIndices", &PdfBase::setIndices)
        .def("SigGenSetIndices", &PdfBase::SigGenSetIndices)
        .def("setCommonNorm", &PdfBase::setCommonNorm, py::arg("v") = true)
        .def("setSeparateNorm", &PdfBase::setSeparateNorm, py::arg("v") = true)
        .def("fillMCDataSimple", &PdfBase::fillMCDataSimple, py::arg("events"), py::arg("seed") = 0)
        .def("fitTo", &PdfBase::fitTo, py::arg("data"), py::arg("verbosity") = 3, py::keep_alive<1, 2>())
        .def("fit", &PdfBase::fit, py::arg("verbosity") = 3)
        .def("__repr__", [](const PdfBase &p) { return repr_of(p); });

    py::class_<GooPdf, PdfBase>(m, "GooPdf")
        .def("setFitControl", &GooPdf::setFitControl, py::arg("fc"))
        .def("integrate", &GooPdf::integrate)
        .def("hasAnalyticIntegral", &GooPdf::hasAnalyticIntegral)
        .def("getValue", &GooPdf::getValue, py::arg("evalfunc") = EvalFunc::Eval)
        .def("evaluateAtPoints", &GooPdf::evaluateAtPoints)
        .def("getCompProbsAtDataPoints", &GooPdf::getCompProbsAtDataPoints)
        .def("evaluateAtData",
             [](GooPdf &p, EvalFunc f) { return to_numpy(p.evaluateAtData(f)); }, py::arg("evalfunc") = EvalFunc::Prob)
        .def("makeGrid", &GooPdf::makeGrid)
        .def("reduce_with_bins", &GooPdf::reduce_with_bins)
        .def("get_bin_grid_size", &GooPdf::get_bin_grid_size)
        .def("get_bin_grid_volume", &GooPdf::get_bin_grid_volume)
        .def("setParameterConstantness", &GooPdf::setParameterConstantness, py::arg("constant") = true)
        .def("scan", [](GooPdf &p, Observable v) {
            std::vector<fptype> vals;
            p.scan(v, vals);
            return vals;
        });

    py::class_<GaussianPdf, GooPdf>(m, "GaussianPdf")
        .def(py::init<std::string, Observable, Variable, Variable>(), py::arg("name"), py::arg("x"), py::arg("mean"),
             py::arg("sigma"));
    py::class_<ExpPdf, GooPdf>(m, "ExpPdf")
        .def(py::init<std::string, Observable, Variable>(), py::arg("name"), py::arg("x"), py::arg("alpha"))
        .def(py::init<std::string, Observable, Variable, Variable>(), py::arg("name"), py::arg("x"), py::arg("alpha"),
             py::arg("offset"))
        .def(py::init([](std::string n, Observable x, std::vector<Variable> w) { return new ExpPdf(n, x, w); }),
             py::arg("name"), py::arg("x"), py::arg("weights"))
        .def(py::init([](std::string n, Observable x, std::vector<Variable> w, Variable off) {
                 return new ExpPdf(n, x, w, off);
             }),
             py::arg("name"), py::arg("x"), py::arg("weights"), py::arg("offset"));
    py::class_<ExpGausPdf, GooPdf>(m, "ExpGausPdf")
        .def(py::init<std::string, Observable, Variable, Variable, Variable>(), py::arg("name"), py::arg("x"),
             py::arg("mean"), py::arg("sigma"), py::arg("tau"));
    py::class_<PolynomialPdf, GooPdf>(m, "PolynomialPdf")
        .def(py::init<std::string, Observable, std::vector<Variable>, unsigned int>(), py::arg("name"), py::arg("x"),
             py::arg("weights"), py::arg("lowestDegree") = 0)
        .def(py::init<std::string, Observable, std::vector<Variable>, Variable, unsigned int>(), py::arg("name"),
             py::arg("x"), py::arg("weights"), py::arg("x0"), py::arg("lowestDegree") = 0)
        .def(py::init<std::string, std::vector<Observable>, std::vector<Variable>, std::vector<Variable>,
                      unsigned int>(),
             py::arg("name"), py::arg("obses"), py::arg("coeffs"), py::arg("offsets"), py::arg("maxDegree") = 0);

    py::class_<CombinePdf, GooPdf>(m, "CombinePdf");
    py::class_<AddPdf, CombinePdf>(m, "AddPdf")
        .def(py::init<std::string, std::vector<Variable>, std::vector<PdfBase *>>(), py::arg("name"),
             py::arg("weights"), py::arg("comps"), py::keep_alive<1, 4>())
        .def(py::init<std::string, Variable, PdfBase *, PdfBase *>(), py::arg("name"), py::arg("frac"),
             py::arg("func1"), py::arg("func2"), py::keep_alive<1, 4>(), py::keep_alive<1, 5>());
    py::class_<ProdPdf, CombinePdf>(m, "ProdPdf")
        .def(py::init<std::string, std::vector<PdfBase *>>(), py::arg("name"), py::arg("comps"),
             py::keep_alive<1, 3>());

    // ------------------------------------------------------------ Dalitz
    py::enum_<DaughterPair>(m, "DaughterPair")
        .value("PAIR_12", PAIR_12)
        .value("PAIR_13", PAIR_13)
        .value("PAIR_23", PAIR_23)
        .export_values();

    py::class_<AmpComponent, GooPdf>(m, "AmpComponent");
    py::class_<ResonancePdf, AmpComponent>(m, "ResonancePdf")
        .def("get_amp_real", &ResonancePdf::get_amp_real)
        .def("get_amp_img", &ResonancePdf::get_amp_img);
    auto mres = m.def_submodule("Resonances");
    py::class_<Resonances::RBW, ResonancePdf>(mres, "RBW")
        .def(py::init<std::string, Variable, Variable, Variable, Variable, unsigned int, unsigned int, bool, bool>(),
             py::arg("name"), py::arg("ar"), py::arg("ai"), py::arg("mass"), py::arg("width"), py::arg("sp"),
             py::arg("cyc"), py::arg("norm") = true, py::arg("sym") = false);
    py::class_<Resonances::GS, ResonancePdf>(mres, "GS")
        .def(py::init<std::string, Variable, Variable, Variable, Variable, unsigned int, unsigned int, bool, bool>(),
             py::arg("name"), py::arg("ar"), py::arg("ai"), py::arg("mass"), py::arg("width"), py::arg("sp"),
             py::arg("cyc"), py::arg("norm") = true, py::arg("sym") = false);
    py::class_<Resonances::NonRes, ResonancePdf>(mres, "NonRes")
        .def(py::init<std::string, Variable, Variable>(), py::arg("name"), py::arg("ar"), py::arg("ai"));

    py::class_<DecayInfo3>(m, "DecayInfo3", py::dynamic_attr())
        .def(py::init<>())
        .def_readwrite("motherMass", &DecayInfo3::motherMass)
        .def_readwrite("daug1Mass", &DecayInfo3::daug1Mass)
        .def_readwrite("daug2Mass", &DecayInfo3::daug2Mass)
        .def_readwrite("daug3Mass", &DecayInfo3::daug3Mass)
        .def_readwrite("meson_radius", &DecayInfo3::meson_radius)
        .def_readwrite("mother_meson_radius", &DecayInfo3::mother_meson_radius)
        .def_property(
            "resonances", [](DecayInfo3 &d) { return d.resonances; },
            [](py::object self, py::list res) {
                DecayInfo3 &d = self.cast<DecayInfo3 &>();
                d.resonances.clear();
                for(auto r : res)
                    d.resonances.push_back(r.cast<ResonancePdf *>());
                self.attr("_resonance_refs") = res;  // keep the Python objects alive
            });

    py::class_<Amp3Body, CombinePdf>(m, "Amp3Body")
        .def(py::init<std::string, Observable, Observable, EventNumber, DecayInfo3, GooPdf *>(), py::arg("name"),
             py::arg("m12"), py::arg("m13"), py::arg("eventNumber"), py::arg("decay"), py::arg("eff"),
             py::keep_alive<1, 6>(), py::keep_alive<1, 7>())
        .def("setDataSize", &Amp3Body::setDataSize, py::arg("dataSize"), py::arg("evtSize") = 3, py::arg("offset") = 0)
        .def("setForceIntegrals", &Amp3Body::setForceIntegrals, py::arg("f") = true)
        .def("getCachedWave", &Amp3Body::getCachedWave)
        .def("sumCachedWave", &Amp3Body::sumCachedWave)
        .def("fit_fractions", &Amp3Body::fit_fractions)
        .def(
            "integralAtPoints",
            [](Amp3Body &p, py::array_t<double, py::array::c_style | py::array::forcecast> m12,
               py::array_t<double, py::array::c_style | py::array::forcecast> m13,
               py::array_t<double, py::array::c_style | py::array::forcecast> w) {
                std::vector<double> a(m12.data(), m12.data() + m12.size());
                std::vector<double> b(m13.data(), m13.data() + m13.size());
                std::vector<double> c(w.data(), w.data() + w.size());
                return p.integralAtPoints(a, b, c);
            },
            py::arg("m12"), py::arg("m13"), py::arg("w"))
        .def("getIntegrals", &Amp3Body::getIntegrals)
        .def("getEfficiencyIndex", &Amp3Body::getEfficiencyIndex)
        .def("setNormalisationPoints", &Amp3Body::setNormalisationPoints, py::arg("n"), py::arg("seed"),
             py::arg("proposal") = nullptr, py::keep_alive<1, 4>())
        .def("setNormalisationGrid", &Amp3Body::setNormalisationGrid)
        .def("refreshCacheOnly", &Amp3Body::refreshCacheOnly)
        .def(
            "setNormalisationRule",
            [](Amp3Body &p, py::array_t<double, py::array::c_style | py::array::forcecast> m12,
               py::array_t<double, py::array::c_style | py::array::forcecast> m13,
               py::array_t<double, py::array::c_style | py::array::forcecast> w) {
                p.setNormalisationRule(std::vector<double>(m12.data(), m12.data() + m12.size()),
                                       std::vector<double>(m13.data(), m13.data() + m13.size()),
                                       std::vector<double>(w.data(), w.data() + w.size()));
            },
            py::arg("m12"), py::arg("m13"), py::arg("w"))
        .def("getNormalisationPointCount", &Amp3Body::getNormalisationPointCount)
        .def("normalisationRelativeVariance", &Amp3Body::normalisationRelativeVariance, py::arg("batches") = 16)
        .def("generateEvents", &Amp3Body::generateEvents, py::arg("out"), py::arg("n"), py::arg("seed"),
             py::arg("fmax") = 0.0)
        .def_static("resetCacheCounter", &Amp3Body::resetCacheCounter);

    py::class_<DecayInfo3t, DecayInfo3>(m, "DecayInfo3t", py::dynamic_attr())
        .def(py::init<Variable, Variable, Variable, Variable, Variable>(), py::arg("tau"), py::arg("xmixing"),
             py::arg("ymixing"), py::arg("deltax"), py::arg("deltay"))
        .def_readonly("_tau", &DecayInfo3t::_tau)
        .def_readonly("_xmixing", &DecayInfo3t::_xmixing)
        .def_readonly("_ymixing", &DecayInfo3t::_ymixing)
        .def_readonly("_deltax", &DecayInfo3t::_deltax)
        .def_readonly("_deltay", &DecayInfo3t::_deltay);

    py::class_<WaveHolder_s>(m, "WaveHolder_s")
        .def_readonly("ai_real", &WaveHolder_s::ai_real)
        .def_readonly("ai_imag", &WaveHolder_s::ai_imag)
        .def_readonly("bi_real", &WaveHolder_s::bi_real)
        .def_readonly("bi_imag", &WaveHolder_s::bi_imag);

    py::class_<MixingTimeResolution, GooPdf>(m, "MixingTimeResolution")
        .def("normalization", &MixingTimeResolution::normalization, py::arg("di1"), py::arg("di2"), py::arg("di3"),
             py::arg("di4"), py::arg("tau"), py::arg("xmixing"), py::arg("ymixing"))
        .def("getDeviceFunction", &MixingTimeResolution::getDeviceFunction)
        .def("getCalcTauIdx", &MixingTimeResolution::getCalcTauIdx)
        .def("setCalcTauIdx", &MixingTimeResolution::setCalcTauIdx);
    py::class_<TruthResolution, MixingTimeResolution>(m, "TruthResolution").def(py::init<>());
    py::class_<ThreeGaussResolution, MixingTimeResolution>(m, "ThreeGaussResolution")
        .def(py::init<Variable, Variable, Variable, Variable, Variable, Variable, Variable, Variable, Variable,
                      Variable, Variable, Variable, Variable>(),
             py::arg("cf"), py::arg("tf"), py::arg("cb"), py::arg("cs"), py::arg("tb"), py::arg("ts"), py::arg("ob"),
             py::arg("os"), py::arg("sb"), py::arg("common_bias_offset"), py::arg("common_width_scale"),
             py::arg("common_bias_offset2"), py::arg("common_width_scale2"));

    py::class_<Amp3Body_TD, CombinePdf>(m, "Amp3Body_TD")
        .def(py::init<std::string, Observable, Observable, Observable, Observable, EventNumber, DecayInfo3t,
                      MixingTimeResolution *, GooPdf *>(),
             py::arg("name"), py::arg("dtime"), py::arg("sigmat"), py::arg("m12"), py::arg("m13"),
             py::arg("eventNumber"), py::arg("decay"), py::arg("r"), py::arg("eff"), py::keep_alive<1, 8>(),
             py::keep_alive<1, 9>(), py::keep_alive<1, 10>())
        .def(py::init<std::string, Observable, Observable, Observable, Observable, EventNumber, DecayInfo3t,
                      MixingTimeResolution *, GooPdf *, Observable *, Observable *>(),
             py::arg("name"), py::arg("dtime"), py::arg("sigmat"), py::arg("m12"), py::arg("m13"),
             py::arg("eventNumber"), py::arg("decay"), py::arg("r"), py::arg("eff"), py::arg("mistag"),
             py::arg("charmtag"), py::keep_alive<1, 8>(), py::keep_alive<1, 9>(), py::keep_alive<1, 10>())
        .def(py::init<std::string, Observable, Observable, Observable, Observable, EventNumber, DecayInfo3t,
                      std::vector<MixingTimeResolution *> &, GooPdf *, Observable, Observable *, Observable *>(),
             py::arg("name"), py::arg("dtime"), py::arg("sigmat"), py::arg("m12"), py::arg("m13"),
             py::arg("eventNumber"), py::arg("decay"), py::arg("r"), py::arg("eff"), py::arg("md0"),
             py::arg("mistag") = nullptr, py::arg("charmtag") = nullptr)
        .def("setDataSize", &Amp3Body_TD::setDataSize, py::arg("dataSize"), py::arg("evtSize") = 5,
             py::arg("offset") = 0)
        .def("setD0Fraction", &Amp3Body_TD::setD0Fraction)
        .def("getD0Fraction", &Amp3Body_TD::getD0Fraction)
        .def("setForceIntegrals", &Amp3Body_TD::setForceIntegrals, py::arg("f") = true)
        .def("getCachedWaveNoCopy", &Amp3Body_TD::getCachedWaveNoCopy)
        .def("getFractions", &Amp3Body_TD::getFractions)
        .def("getIntegralsAA", &Amp3Body_TD::getIntegralsAA)
        .def("getIntegralsAB", &Amp3Body_TD::getIntegralsAB)
        .def("getIntegralsBB", &Amp3Body_TD::getIntegralsBB)
        .def("getDecayInfo", &Amp3Body_TD::getDecayInfo, py::return_value_policy::reference_internal)
        .def_static("resetCacheCounter", &Amp3Body_TD::resetCacheCounter);

    py::class_<DalitzProposal>(m, "DalitzProposal")
        .def(py::init<Observable, Observable, unsigned long long, int>(), py::arg("m12"), py::arg("m13"),
             py::arg("seed") = 1, py::arg("batch_cap") = 65536)
        .def("train", &DalitzProposal::train, py::arg("events"), py::arg("steps") = 2000, py::arg("batch") = 8192,
             py::arg("lr") = 3e-3f, py::arg("seed") = 1)
        .def("trainOnModel", &DalitzProposal::trainOnModel, py::arg("model"), py::arg("nevents"),
             py::arg("gen_seed"), py::arg("steps") = 2000, py::arg("batch") = 8192, py::arg("lr") = 3e-3f,
             py::arg("seed") = 1)
        .def("build", &DalitzProposal::build, py::arg("grid") = 512, py::arg("alpha") = 0.1)
        .def("isBuilt", &DalitzProposal::isBuilt)
        .def("evaluate",
             [](const DalitzProposal &q, py::array_t<float, py::array::c_style | py::array::forcecast> xy) {
                 std::vector<float> v(xy.data(), xy.data() + xy.size());
                 auto out = q.evaluate(v);
                 py::array_t<float> a(out.size());
                 std::copy(out.begin(), out.end(), a.mutable_data());
                 return a;
             })
        .def("getParameters", &DalitzProposal::getParameters)
        .def("setParameters", &DalitzProposal::setParameters)
        .def("setReferenceVariance", &DalitzProposal::setReferenceVariance)
        .def("getReferenceVariance", &DalitzProposal::getReferenceVariance)
        .def("needsRetraining", &DalitzProposal::needsRetraining, py::arg("relvar"), py::arg("factor") = 4.0)
        .def(
            "setRetraining",
            [](DalitzProposal &q, int64_t events, int steps, int batch, float lr, int grid, double alpha,
               double factor) {
                DalitzProposal::Retraining r;
                r.events = events;
                r.steps  = steps;
                r.batch  = batch;
                r.lr     = lr;
                r.grid   = grid;
                r.alpha  = alpha;
                r.factor = factor;
                q.setRetraining(r);
            },
            py::arg("events") = 2000000, py::arg("steps") = 500, py::arg("batch") = 8192, py::arg("lr") = 6e-3f,
            py::arg("grid") = 512, py::arg("alpha") = 0.1, py::arg("factor") = 4.0)
        .def("disableRetraining", &DalitzProposal::disableRetraining)
        .def("retrainingEnabled", &DalitzProposal::retrainingEnabled)
        .def("retrainCount", &DalitzProposal::retrainCount);

    m.def("inDalitz", &inDalitz);

    // ------------------------------------------------------------ fitting
    py::class_<M2::FunctionMinimum>(m, "FunctionMinimum")
        .def("IsValid", &M2::FunctionMinimum::IsValid)
        .def("UserState", &M2::FunctionMinimum::UserState, py::return_value_policy::reference_internal)
        .def("UserParameters", &M2::FunctionMinimum::UserParameters, py::return_value_policy::reference_internal)
        .def("UserCovariance", &M2::FunctionMinimum::UserCovariance)
        .def("Up", &M2::FunctionMinimum::Up)
        .def("HasValidParameters", &M2::FunctionMinimum::HasValidParameters)
        .def("HasValidCovariance", &M2::FunctionMinimum::HasValidCovariance)
        .def("HasPosDefCovar", &M2::FunctionMinimum::HasPosDefCovar)
        .def("HasMadePosDefCovar", &M2::FunctionMinimum::HasMadePosDefCovar)
        .def("HesseFailed", &M2::FunctionMinimum::HesseFailed)
        .def("HasCovariance", &M2::FunctionMinimum::HasCovariance)
        .def("HasAccurateCovar", &M2::FunctionMinimum::HasAccurateCovar)
        .def("IsAboveMaxEdm", &M2::FunctionMinimum::IsAboveMaxEdm)
        .def("HasReachedCallLimit", &M2::FunctionMinimum::HasReachedCallLimit)
        .def("Fval", &M2::FunctionMinimum::Fval)
        .def("Edm", &M2::FunctionMinimum::Edm)
        .def("NFcn", &M2::FunctionMinimum::NFcn)
        .def("values", [](const M2::FunctionMinimum &f) { return f.UserParameters().Params(); })
        .def("errors", [](const M2::FunctionMinimum &f) { return f.UserParameters().Errors(); })
        .def("names",
             [](const M2::FunctionMinimum &f) {
                 std::vector<std::string> n;
                 for(auto &p : f.UserParameters().Parameters())
                     n.push_back(p.name);
                 return n;
             })
        .def("covariance", [](const M2::FunctionMinimum &f) { return f.UserState().cov_; })
        .def("__repr__", [](const M2::FunctionMinimum &f) { return repr_of(f); });

    // python/goofit/FitManager.cpp:19-30 of the reference
    py::class_<Params>(m, "Params")
        .def(py::init<PdfBase &>(), py::arg("pdf"), py::keep_alive<1, 2>())
        .def("SetGooFitParams", &Params::SetGooFitParams, py::arg("input"))
        .def("size", &Params::size)
        .def("make_minuit_vector", &Params::make_minuit_vector)
        .def("from_minuit_vector", &Params::from_minuit_vector, py::arg("values"), py::arg("force_changed") = false)
        .def_property("record", &Params::get_recorded,
                      [](Params &p, bool r) { p.set_record(r); })
        .def_property_readonly("recorded", &Params::get_recorded);
    py::class_<FCN>(m, "FCN")
        .def(py::init<Params &>(), py::arg("params"), py::keep_alive<1, 2>())
        .def("__call__", [](const FCN &f, const std::vector<double> &x) { return f(x); })
        .def("Up", &FCN::Up);

    py::class_<BatchFitter>(m, "BatchFitter")
        .def(py::init<GooPdf *, const std::vector<int64_t> &, const std::vector<int64_t> &>(), py::arg("pdf"),
             py::arg("first"), py::arg("count"), py::keep_alive<1, 2>())
        .def("fit", &BatchFitter::fit, py::arg("threads") = 256, py::arg("strategy") = 1, py::arg("tolerance") = 0.1,
             py::arg("autotune") = true, py::call_guard<py::gil_scoped_release>())
        .def("evaluate", &BatchFitter::evaluate)
        .def("useGraph", &BatchFitter::useGraph)
        .def("useMultiPoint", &BatchFitter::useMultiPoint)
        .def("multiPoint", &BatchFitter::multiPoint)
        .def("useDeviceNormalisation", &BatchFitter::useDeviceNormalisation)
        .def("setStages", &BatchFitter::setStages)
        .def("__len__", &BatchFitter::size)
        .def("values", &BatchFitter::values)
        .def("errors", &BatchFitter::errors)
        .def("valid", &BatchFitter::valid)
        .def("nfcn", &BatchFitter::nfcn)
        .def("fval", &BatchFitter::fval)
        .def("stats", [](const BatchFitter &b) {
            const BatchFitStats &s = b.stats();
            py::dict d;
            d["launches"]   = s.launches;
            d["fcn_calls"]  = s.fcn_calls;
            d["seconds"]    = s.seconds;
            d["gpu_wait_s"] = s.gpu_wait_s;
            d["host_prep_s"] = s.host_prep_s;
            d["gather_s"]   = s.gather_s;
            d["chunk"]      = s.chunk;
            d["device_norms"] = s.device_norms;
            d["two_stages"]   = s.two_stages;
            return d;
        });

    py::class_<FitManagerMinuit2>(m, "FitManager")
        .def(py::init<PdfBase *>(), py::arg("pdf"), py::keep_alive<1, 2>())
        .def("fit", &FitManagerMinuit2::fit)
        .def("setMaxCalls", &FitManagerMinuit2::setMaxCalls, py::arg("max_calls") = 0)
        .def("setVerbosity", &FitManagerMinuit2::setVerbosity)
        .def("getVerbosity", &FitManagerMinuit2::getVerbosity)
        .def("setMinos", &FitManagerMinuit2::setMinos, py::arg("minos_flag") = true)
        .def("getMinosErrors", &FitManagerMinuit2::getMinosErrors)
        .def("setStrategy", &FitManagerMinuit2::setStrategy)
        .def("setTolerance", &FitManagerMinuit2::setTolerance)
        .def("setBatchedDerivatives", &FitManagerMinuit2::setBatchedDerivatives, py::arg("on") = true)
        .def("getBatchedDerivatives", &FitManagerMinuit2::getBatchedDerivatives)
        .def_property("verbosity", &FitManagerMinuit2::getVerbosity, &FitManagerMinuit2::setVerbosity)
        .def_property("max_calls", nullptr, [](FitManagerMinuit2 &f, unsigned v) { f.setMaxCalls(v); })
        .def_property_readonly("params", &FitManagerMinuit2::getParams, py::return_value_policy::reference_internal)
        .def_property_readonly("fcn", &FitManagerMinuit2::getFCN, py::return_value_policy::reference_internal)
        .def("__nonzero__", [](const FitManagerMinuit2 &f) { return (bool)f; })
        .def("getLastFitTime", &FitManagerMinuit2::getLastFitTime)
        .def("getMnScan", &FitManagerMinuit2::getMnScan, py::keep_alive<0, 1>())
        .def("minuit_vector", [](FitManagerMinuit2 &f) { return f.getParams()->make_minuit_vector(); })
        .def("free_mask",
             [](FitManagerMinuit2 &f) {
                 std::vector<bool> m;
                 for(auto &p : f.getParams()->Parameters())
                     m.push_back(!p.fixed);
                 return m;
             })
        .def("evaluate", [](FitManagerMinuit2 &f, const std::vector<double> &x) { return (*f.getFCN())(x); },
             "One FCN call: push the Minuit vector into the PDF and return its NLL")
        .def(
            "time_calls",
            [](FitManagerMinuit2 &f, const std::vector<std::vector<double>> &xs) {
                // one FCN call per vector, in order, from C++ as the minimiser makes them
                const auto &fcn = *f.getFCN();
                std::vector<double> out(xs.size());
                const auto t0 = std::chrono::steady_clock::now();
                for(size_t k = 0; k < xs.size(); k++)
                    out[k] = fcn(xs[k]);
                const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
                return py::make_tuple(s, out);
            },
            py::arg("xs"), "Seconds for one FCN call per Minuit vector, made one after the other, and the NLLs")
        .def("__bool__", [](const FitManagerMinuit2 &f) { return (bool)f; })
        .def("__int__", [](const FitManagerMinuit2 &f) { return (int)f; });

    // ------------------------------------------------------------ Minuit2 (python/Minuit2/*.cpp of the reference)
    {
        auto mm = m.def_submodule("Minuit2", "The Minuit2 API used by GooFit, on the goofit_amd minimiser");
        struct PyFCN : M2::FCNBase {
            using M2::FCNBase::FCNBase;
            double operator()(const std::vector<double> &v) const override {
                PYBIND11_OVERRIDE_PURE_NAME(double, M2::FCNBase, "__call__", operator(), v);
            }
            double Up() const override { PYBIND11_OVERRIDE_PURE(double, M2::FCNBase, Up, ); }
            double ErrorDef() const override { PYBIND11_OVERRIDE(double, M2::FCNBase, ErrorDef, ); }
        };
        py::class_<M2::FCNBase, PyFCN>(mm, "FCNBase")
            .def(py::init<>())
            .def("__call__", &M2::FCNBase::operator())
            .def("ErrorDef", &M2::FCNBase::ErrorDef)
            .def("Up", &M2::FCNBase::Up);
        py::class_<M2::MinuitParameter>(mm, "MinuitParameter")
            .def("Name", [](const M2::MinuitParameter &p) { return p.name; })
            .def("GetName", [](const M2::MinuitParameter &p) { return p.name; })
            .def("Value", [](const M2::MinuitParameter &p) { return p.value; })
            .def("Error", [](const M2::MinuitParameter &p) { return p.error; })
            .def("IsFixed", [](const M2::MinuitParameter &p) { return p.fixed; })
            .def("IsConst", [](const M2::MinuitParameter &p) { return p.fixed && p.error == 0; })
            .def("HasLimits", [](const M2::MinuitParameter &p) { return p.has_lower && p.has_upper; })
            .def("HasLowerLimit", [](const M2::MinuitParameter &p) { return p.has_lower; })
            .def("HasUpperLimit", [](const M2::MinuitParameter &p) { return p.has_upper; })
            .def("LowerLimit", [](const M2::MinuitParameter &p) { return p.lower; })
            .def("UpperLimit", [](const M2::MinuitParameter &p) { return p.upper; });
        py::class_<M2::MnUserParameters>(mm, "MnUserParameters")
            .def(py::init<>())
            .def("Add", (bool(M2::MnUserParameters::*)(const std::string &, double)) & M2::MnUserParameters::Add)
            .def("Add",
                 (bool(M2::MnUserParameters::*)(const std::string &, double, double)) & M2::MnUserParameters::Add)
            .def("Add", (bool(M2::MnUserParameters::*)(const std::string &, double, double, double, double))
                            & M2::MnUserParameters::Add)
            .def("Parameter", [](const M2::MnUserParameters &u, unsigned int i) { return u.Parameters().at(i); })
            .def("Parameters", &M2::MnUserParameters::Parameters)
            .def("Params", &M2::MnUserParameters::Params)
            .def("Errors", &M2::MnUserParameters::Errors)
            .def("Value", &M2::MnUserParameters::Value)
            .def("Error", &M2::MnUserParameters::Error)
            .def("SetValue", &M2::MnUserParameters::SetValue)
            .def("SetError", &M2::MnUserParameters::SetError)
            .def("SetLimits", &M2::MnUserParameters::SetLimits)
            .def("SetLowerLimit", &M2::MnUserParameters::SetLowerLimit)
            .def("SetUpperLimit", &M2::MnUserParameters::SetUpperLimit)
            .def("RemoveLimits", &M2::MnUserParameters::RemoveLimits)
            .def("Fix", &M2::MnUserParameters::Fix)
            .def("Release", &M2::MnUserParameters::Release)
            .def("Index", &M2::MnUserParameters::Index)
            .def("GetName", &M2::MnUserParameters::GetName)
            .def("NParameters", &M2::MnUserParameters::NParameters)
            .def("VariableParameters", &M2::MnUserParameters::VariableParameters);
        py::class_<M2::MnUserParameterState>(mm, "MnUserParameterState")
            .def("Value", (double(M2::MnUserParameterState::*)(unsigned int) const) & M2::MnUserParameterState::Value,
                 py::arg("n"))
            .def("Error", (double(M2::MnUserParameterState::*)(unsigned int) const) & M2::MnUserParameterState::Error,
                 py::arg("n"))
            .def("Covariance", &M2::MnUserParameterState::Covariance)
            .def("Fval", &M2::MnUserParameterState::Fval)
            .def("Edm", &M2::MnUserParameterState::Edm)
            .def("NFcn", &M2::MnUserParameterState::NFcn)
            .def("IsValid", &M2::MnUserParameterState::IsValid);
        py::class_<M2::MnUserCovariance>(mm, "MnUserCovariance")
            .def("Nrow", &M2::MnUserCovariance::Nrow)
            .def("size", &M2::MnUserCovariance::size)
            .def("__call__", &M2::MnUserCovariance::operator())
            .def("to_matrix", [](const M2::MnUserCovariance &c) {
                py::array_t<double> a({(py::ssize_t)c.Nrow(), (py::ssize_t)c.Nrow()});
                std::copy(c.Data().begin(), c.Data().end(), a.mutable_data());
                return a;
            });
        py::class_<M2::MnMigrad>(mm, "MnMigrad")
            .def(py::init([](const M2::FCNBase &fcn, const M2::MnUserParameters &par, unsigned int stra) {
                     return new M2::MnMigrad(fcn, par, stra);
                 }),
                 py::arg("fcn"), py::arg("par"), py::arg("stra") = 1, py::keep_alive<1, 2>())
            .def("__call__", &M2::MnMigrad::operator(), py::arg("maxfcn") = 0, py::arg("tolerance") = 0.1);
        py::class_<M2::MnMinos>(mm, "MnMinos")
            .def(py::init<const M2::FCNBase &, const M2::FunctionMinimum &, unsigned int>(), py::arg("fcn"),
                 py::arg("min"), py::arg("stra") = 1, py::keep_alive<1, 2>(), py::keep_alive<1, 3>())
            .def("__call__", &M2::MnMinos::operator(), py::arg("par"), py::arg("maxcalls") = 0,
                 "(lower, upper) error of parameter par")
            .def("Lower", &M2::MnMinos::Lower, py::arg("par"), py::arg("maxcalls") = 0)
            .def("Upper", &M2::MnMinos::Upper, py::arg("par"), py::arg("maxcalls") = 0);
        py::class_<M2::MnScan>(mm, "MnScan")
            .def(py::init<const M2::FCNBase &, const M2::MnUserParameters &, unsigned int>(), py::arg("fcn"),
                 py::arg("par"), py::arg("stra") = 1, py::keep_alive<1, 2>())
            .def("Scan", &M2::MnScan::Scan, "Scan parameter.", py::arg("par"), py::arg("maxsteps") = 41,
                 py::arg("low") = 0.0, py::arg("high") = 0.0);
        mm.attr("FunctionMinimum") = m.attr("FunctionMinimum");
        auto mp = mm.def_submodule("MnPrint");
        mp.def("GlobalLevel", &M2::MnPrint::GlobalLevel);
        mp.def("SetGlobalLevel", &M2::MnPrint::SetGlobalLevel);
    }

    // ------------------------------------------------------------ engine state
    auto me = m.def_submodule("engine");
    me.def("set_device", &engine::set_device);
    me.def("get_device", &engine::get_device);
    me.def("release", &engine::release);
    me.def("flush", &engine::flush);
    // launch-shape knobs of this thread's context (gf_tune codes)
    me.def("set_tuning", [](int key, int value) {
        engine::check(gf_ctx_set_tuning(engine::ctx(), key, value), "gf_ctx_set_tuning");
    });
    me.def("get_tuning", [](int key) {
        int v = 0;
        engine::check(gf_ctx_get_tuning(engine::ctx(), key, &v), "gf_ctx_get_tuning");
        return v;
    });
    me.attr("TUNE_INLINE_ARGS")         = (int)GF_TUNE_INLINE_ARGS;
    me.attr("TUNE_FAST_DALITZ")         = (int)GF_TUNE_FAST_DALITZ;
    me.attr("TUNE_BLOCKS_PER_CU")       = (int)GF_TUNE_BLOCKS_PER_CU;
    me.attr("TUNE_FAST_BLOCKS_PER_CU")  = (int)GF_TUNE_FAST_BLOCKS_PER_CU;
    me.attr("TUNE_SPIN_US")             = (int)GF_TUNE_SPIN_US;
    me.attr("TUNE_ZERO_OUTSIDE")        = (int)GF_TUNE_ZERO_OUTSIDE;
    me.attr("TUNE_FLAT")                = (int)GF_TUNE_FLAT;
    me.attr("TUNE_FUSED_FINISH")        = (int)GF_TUNE_FUSED_FINISH;
    me.attr("TUNE_COH_MFMA")            = (int)GF_TUNE_COH_MFMA;
    me.attr("TUNE_BATCH_MX_PTS")        = (int)GF_TUNE_BATCH_MX_PTS;
    me.def("callnumber", []() { return host_callnumber; });
    me.def("flattened", []() {
        return py::make_tuple(host_function_table, host_parameters, host_constants, host_observables,
                              host_normalizations);
    });
    me.def("function_names", []() { return host_function_name; });
    // device time of back-to-back launches on the engine's stream (HIP events), ms
    me.def("time_reduce_sweep", [](int metric, int reps) {
        engine::flush();
        double sweep = 0, total = 0;
        engine::check(gf_time_reduce_sweep(engine::ctx(), metric, reps, &sweep, &total), "gf_time_reduce_sweep");
        return py::make_tuple(sweep, total);
    });
    me.def("time_sweep_back_to_back", [](int metric, int reps) {
        engine::flush();
        double ms = 0;
        engine::check(gf_time_sweep_back_to_back(engine::ctx(), metric, reps, &ms), "gf_time_sweep_back_to_back");
        return ms;
    });
    me.def("time_reduce_metric", [](int metric, int reps) {
        engine::flush();
        double ms = 0;
        engine::check(gf_time_reduce_metric(engine::ctx(), metric, reps, &ms), "gf_time_reduce_metric");
        return ms;
    });
    me.def("time_dalitz_cache", [](const Amp3Body &p, int reps) {
        engine::flush();
        double ms = 0;
        engine::check(gf_time_dalitz_cache(engine::ctx(), (int)p.getFunctionIndex(), reps, &ms),
                      "gf_time_dalitz_cache");
        return ms;
    });
    me.def("time_dalitz_gram", [](const Amp3Body &p, std::vector<double> grid6, int reps) {
        engine::flush();
        double ms = 0;
        engine::check(gf_time_dalitz_gram(engine::ctx(), (int)p.getFunctionIndex(), grid6.data(),
                                          (int)p.getEfficiencyIndex(), reps, &ms),
                      "gf_time_dalitz_gram");
        return ms;
    });
}
