// CombinePdfs.cpp -- AddPdf and ProdPdf (src/PDFs/combine/{AddPdf,ProdPdf}.cu of
// the reference): construction and the host side of normalize()/calculateNLL().
#include <goofit/Error.h>
#include <goofit/PDFs/combine/AddPdf.h>
#include <goofit/PDFs/combine/ProdPdf.h>

#include <algorithm>
#include <cmath>

namespace GooFit {

// ---------------------------------------------------------------- AddPdf

AddPdf::AddPdf(std::string n, std::vector<Variable> weights, std::vector<PdfBase *> comps)
    : CombinePdf("AddPdf", n)
    , extended(true) {
    if(weights.size() != comps.size() && (weights.size() + 1) != comps.size())
        throw GeneralError("Size of weights {} (+1) != comps {}", weights.size(), comps.size());
    for(PdfBase *p : comps) {
        if(p == nullptr)
            throw GeneralError("Invalid component");
        components.push_back(p);
    }
    observablesList = getObservables();
    for(auto &w : weights)
        registerParameter(w);
    if(weights.size() < components.size())
        extended = false;
    if(extended)
        registerFunction("ptr_to_AddPdfsExt", GF_FN_ADD_EXT);
    else
        registerFunction("ptr_to_AddPdfs", GF_FN_ADD);
    initialize();
}

AddPdf::AddPdf(std::string n, Variable frac1, PdfBase *func1, PdfBase *func2)
    : CombinePdf("AddPdf", n, frac1)
    , extended(false) {
    if(!func1 || !func2)
        throw GeneralError("Invalid component");
    components.push_back(func1);
    components.push_back(func2);
    observablesList = getObservables();
    registerFunction("ptr_to_AddPdfs", GF_FN_ADD);
    initialize();
}

// AddPdf::normalize, src/PDFs/combine/AddPdf.cu:143-185: components normalise
// separately; the sum itself keeps normalisation 1.
fptype AddPdf::normalize() {
    fptype ret         = 0;
    fptype totalWeight = 0;
    for(size_t i = 0; i + 1 < components.size(); ++i) {
        fptype weight = parametersList[i].getValue();
        totalWeight += weight;
        fptype curr = components[i]->normalize();
        ret += curr * weight;
    }
    fptype last = components.back()->normalize();
    if(extended) {
        fptype lastWeight = parametersList[components.size() - 1].getValue();
        totalWeight += lastWeight;
        ret += last * lastWeight;
        ret /= totalWeight;
    } else {
        ret += (1 - totalWeight) * last;
    }
    setNormalization(1.0);
    if(getCommonNorm())
        for(auto *c : components)
            c->setNormalization(1.0 / ret);
    return ret;
}

// The sum keeps normalisation 1 and its components normalise separately, unless they
// share the common norm 1 / sum (which moves with every weight: no device op).
bool AddPdf::batchNormOps(std::vector<gf_norm_op> &ops, std::vector<double> &table) const {
    if(getCommonNorm())
        return GooPdf::batchNormOps(ops, table);
    for(PdfBase *c : components) {
        const auto *g = dynamic_cast<const GooPdf *>(c);
        if(!g || !g->batchNormOps(ops, table))
            return false;
    }
    return true;
}

// The weights of an extended sum are the yields (nllExtraTermFor adds them in this order).
std::vector<unsigned int> AddPdf::extendedYieldSlots() const {
    std::vector<unsigned int> out;
    if(extended)
        for(size_t i = 0; i < components.size(); ++i)
            out.push_back(parametersIdx + 1 + (unsigned int)i);
    return out;
}

// AddPdf::calculateNLL, AddPdf.cu:187-204: extended likelihood adds E - N log E.
double AddPdf::calculateNLL() {
    double ret = GooPdf::calculateNLL();
    ret += nllExtraTerm();
    return ret;
}

double AddPdf::nllExtraTerm() const { return nllExtraTermFor(numEvents); }

double AddPdf::nllExtraTermFor(fptype nevents) const {
    if(!extended)
        return 0.0;
    fptype expEvents = 0;
    for(size_t i = 0; i < components.size(); ++i)
        expEvents += parametersList[i].getValue();
    return expEvents - nevents * std::log(expEvents);
}

// ---------------------------------------------------------------- ProdPdf

ProdPdf::ProdPdf(std::string n, std::vector<PdfBase *> comps)
    : CombinePdf("ProdPdf", n)
    , varOverlaps(false) {
    for(PdfBase *p : comps)
        components.push_back(p);
    observablesList = getObservables();
    registerConstant((fptype)components.size());
    // components sharing an observable cannot be normalised one by one
    // (ProdPdf.cu:44-68: each component is compared with the one before it)
    std::vector<Observable> observableCheck;
    for(PdfBase *p : comps) {
        if(varOverlaps)
            continue;
        std::vector<Observable> curr = p->getObservables();
        for(Observable &o : curr) {
            if(std::find(observableCheck.begin(), observableCheck.end(), o) == observableCheck.end())
                continue;
            varOverlaps = true;
            break;
        }
        observableCheck = p->getObservables();
    }
    if(varOverlaps)
        for(PdfBase *p : comps)
            if(p->getSeparateNorm())
                varOverlaps = false;
    registerFunction("ptr_to_ProdPdfs", GF_FN_PROD);
    initialize();
}

// Separately normalised components: their ops (the product keeps normalisation 1);
// overlapping observables normalise the product numerically.
bool ProdPdf::batchNormOps(std::vector<gf_norm_op> &ops, std::vector<double> &table) const {
    if(varOverlaps)
        return GooPdf::batchNormOps(ops, table);
    for(PdfBase *c : components) {
        const auto *g = dynamic_cast<const GooPdf *>(c);
        if(!g || !g->batchNormOps(ops, table))
            return false;
    }
    return true;
}

// ProdPdf::normalize, ProdPdf.cu:75-99
fptype ProdPdf::normalize() {
    if(varOverlaps) {
        recursiveSetNormalization(1.0);
        return GooPdf::normalize();
    }
    for(PdfBase *c : components)
        c->normalize();
    setNormalization(1.0);
    return 1.0;
}

}  // namespace GooFit
