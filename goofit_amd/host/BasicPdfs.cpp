// BasicPdfs.cpp -- Gaussian, Exp, ExpGaus and Polynomial PDFs: parameter
// registration and host-side analytic integrals (src/PDFs/basic/*.cu of the
// reference; the per-event functions are in goofit_amd/csrc/gf_device.h).
#include <goofit/Error.h>
#include <goofit/PDFs/basic/ExpGausPdf.h>
#include <goofit/PDFs/basic/ExpPdf.h>
#include <goofit/PDFs/basic/GaussianPdf.h>
#include <goofit/PDFs/basic/PolynomialPdf.h>

#include <cmath>
#include <cstdio>
#include <iostream>

namespace GooFit {

// ---------------------------------------------------------------- Gaussian

GaussianPdf::GaussianPdf(std::string n, Observable _x, Variable mean, Variable sigma)
    : GooPdf("GaussianPdf", n, _x, mean, sigma) {
    registerFunction("ptr_to_Gaussian", GF_FN_GAUSSIAN);
    initialize();
}

// Over all of R, whatever the observable range (GaussianPdf::integrate,
// src/PDFs/basic/GaussianPdf.cu:28-35).
namespace {
fptype gaussian_norm_factor() {
    static const fptype rootPi = std::sqrt(std::atan2(0.0, -1.0));
    return root2 * rootPi;
}
}  // namespace

fptype GaussianPdf::integrate(fptype, fptype) const {
    fptype sigma = host_parameters.at(parametersIdx + 2);
    sigma *= gaussian_norm_factor();
    return sigma;
}

// normalize(): 1 / (sigma * sqrt 2 sqrt pi), the same product on the device
bool GaussianPdf::batchNormOps(std::vector<gf_norm_op> &ops, std::vector<double> &) const {
    if(fitControl && !fitControl->metricIsPdf())
        return true;
    gf_norm_op op{};
    op.kind = GF_NORM_SCALE;
    op.out  = (int32_t)normalIdx + 1;
    op.p    = (int32_t)parametersIdx + 2;
    op.c0   = gaussian_norm_factor();
    ops.push_back(op);
    return true;
}

// ---------------------------------------------------------------- Exp

ExpPdf::ExpPdf(std::string n, Observable _x, Variable alpha)
    : GooPdf("ExpPdf", n, _x, alpha) {
    registerFunction("ptr_to_Exp", GF_FN_EXP);
    initialize();
}

ExpPdf::ExpPdf(std::string n, Observable _x, Variable alpha, Variable offset)
    : GooPdf("ExpPdf", n, _x, offset, alpha) {
    registerFunction("ptr_to_ExpOffset", GF_FN_EXP_OFFSET);
    initialize();
}

ExpPdf::ExpPdf(std::string n, Observable _x, std::vector<Variable> &weights)
    : GooPdf("ExpPdf", n, _x) {
    if(weights.empty())
        throw GeneralError("Weights are empty!");
    for(Variable &w : weights)
        registerParameter(w);
    registerFunction("ptr_to_ExpPoly", GF_FN_EXP_POLY);
    initialize();
}

ExpPdf::ExpPdf(std::string n, Observable _x, std::vector<Variable> &weights, Variable offset)
    : GooPdf("ExpPdf", n, _x, offset) {
    if(weights.empty())
        throw GeneralError("Weights are empty!");
    for(Variable &w : weights)
        registerParameter(w);
    registerFunction("ptr_to_ExpPolyOffset", GF_FN_EXP_POLY_OFFSET);
    initialize();
}

bool ExpPdf::hasAnalyticIntegral() const { return 1 == host_parameters.at(parametersIdx); }

// (e^(alpha hi) - e^(alpha lo)) / alpha, or hi - lo at alpha = 0 (ExpPdf.cu:121-132)
// normalize() of the one-parameter form: 1 / ((exp(alpha hi) - exp(alpha lo)) / alpha)
bool ExpPdf::batchNormOps(std::vector<gf_norm_op> &ops, std::vector<double> &table) const {
    if(fitControl && !fitControl->metricIsPdf())
        return true;
    if(!hasAnalyticIntegral())
        return GooPdf::batchNormOps(ops, table);
    gf_norm_op op{};
    op.kind = GF_NORM_EXP;
    op.out  = (int32_t)normalIdx + 1;
    op.p    = (int32_t)parametersIdx + 1;
    op.c0   = observablesList.at(0).getLowerLimit();
    op.c1   = observablesList.at(0).getUpperLimit();
    ops.push_back(op);
    return true;
}

fptype ExpPdf::integrate(fptype lo, fptype hi) const {
    fptype alpha = host_parameters.at(parametersIdx + 1);
    if(0 == alpha)
        return (hi - lo);
    fptype ret = std::exp(alpha * hi) - std::exp(alpha * lo);
    ret /= alpha;
    return ret;
}

// ---------------------------------------------------------------- ExpGaus

ExpGausPdf::ExpGausPdf(std::string n, Observable _x, Variable mean, Variable sigma, Variable tau)
    : GooPdf("ExpGausPdf", n, _x, mean, sigma, tau) {
    registerFunction("ptr_to_ExpGaus", GF_FN_EXPGAUS);
    initialize();
}

// ---------------------------------------------------------------- Polynomial

PolynomialPdf::PolynomialPdf(std::string n, Observable _x, std::vector<Variable> weights, unsigned int lowestDegree)
    : GooPdf("PolynomialPdf", n, _x) {
    registerConstant(lowestDegree);
    for(Variable &v : weights)
        registerParameter(v);
    registerFunction("ptr_to_Polynomial", GF_FN_POLYNOMIAL);
    initialize();
}

PolynomialPdf::PolynomialPdf(std::string n, Observable _x, std::vector<Variable> weights, Variable x0,
                             unsigned int lowestDegree)
    : GooPdf("PolynomialPdf", n, _x)
    , center(new Variable(x0)) {
    registerConstant(lowestDegree);
    for(Variable &v : weights)
        registerParameter(v);
    registerParameter(x0);
    registerFunction("ptr_to_OffsetPolynomial", GF_FN_OFFSET_POLYNOMIAL);
    initialize();
}

// Multivariate: constants [maxDegree, maxDegree+1, maxDegree+2, ...], parameters
// [offsets..., coefficients...] (PolynomialPdf.cu:166-215); missing coefficients
// and offsets are created fixed at zero.
PolynomialPdf::PolynomialPdf(std::string n, std::vector<Observable> obses, std::vector<Variable> coeffs,
                             std::vector<Variable> offsets, unsigned int maxDegree)
    : GooPdf("PolynomialPdf", n) {
    unsigned int numParameters = 1;
    registerConstant(maxDegree);
    for(unsigned int i = 0; i < obses.size(); ++i) {
        registerObservable(obses[i]);
        numParameters *= (maxDegree + 1 + i);
        registerConstant(maxDegree + 1 + i);
    }
    for(int i = (int)observablesList.size(); i > 1; --i)
        numParameters /= i;
    while(numParameters > coeffs.size()) {
        char name[160];
        snprintf(name, sizeof(name), "%s_extra_coeff_%i", getName().c_str(), (int)coeffs.size());
        coeffs.emplace_back(name, 0);
        std::cout << "Warning: " << getName() << " created dummy variable " << name
                  << " (fixed at zero) to account for all terms.\n";
    }
    while(offsets.size() < obses.size()) {
        char name[160];
        snprintf(name, sizeof(name), "%s_extra_offset_%i", getName().c_str(), (int)offsets.size());
        offsets.emplace_back(name, 0);
    }
    for(auto &o : offsets)
        registerParameter(o);
    for(auto &c : coeffs)
        registerParameter(c);
    registerFunction("ptr_to_MultiPolynomial", GF_FN_MULTI_POLYNOMIAL);
    initialize();
}

// PolynomialPdf::integrate, src/PDFs/basic/PolynomialPdf.cu:210-232 (kept as in the
// reference; the PDF normalises numerically since it reports no analytic integral)
fptype PolynomialPdf::integrate(fptype lo, fptype hi) const {
    fptype lowestDegree = host_constants.at(constantsIdx + 1);
    if(center) {
        hi -= host_observables.at(observablesIdx + 1);
        lo -= host_observables.at(observablesIdx + 2);
    }
    fptype ret = 0;
    for(int i = 2; i < host_parameters.at(parametersIdx) + (center ? 0 : 1); ++i) {
        fptype powerPlusOne = lowestDegree + i - 2;
        fptype curr         = std::pow(hi, powerPlusOne);
        curr -= std::pow(lo, powerPlusOne);
        curr /= powerPlusOne;
        ret += host_parameters.at(parametersIdx + i) * curr;
    }
    return ret;
}

fptype PolynomialPdf::getCoefficient(int coef) {
    if(1 != observablesList.size()) {
        std::cout << "Warning: getCoefficient method of PolynomialPdf not implemented for multi-dimensional "
                     "polynomials. Returning zero, which is very likely wrong.\n";
        return 0;
    }
    if(coef < host_constants.at(constantsIdx + 1))
        return 0;
    if(coef > host_constants.at(constantsIdx + 1) + (host_parameters.at(parametersIdx) - 1))
        return 0;
    fptype norm = normalize();
    norm        = (1.0 / norm);
    fptype param = host_parameters.at(parametersIdx + 2 + coef - int(host_constants.at(constantsIdx + 1)));
    return norm * param;
}

}  // namespace GooFit
