// Amp3Body.cpp -- the three-body Dalitz PDF and its resonances
// (src/PDFs/physics/Amp3Body.cu, resonances/{RBW,GS,NonRes}.cu and
// DalitzPlotHelpers.cu of the reference): tree layout, wave-cache management and
// the host assembly of the normalisation.
#include <goofit/Error.h>
#include <goofit/PDFs/physics/Amp3Body.h>
#include <goofit/PDFs/physics/DalitzProposal.h>
#include <goofit/UnbinnedDataSet.h>
#include <goofit/PDFs/physics/resonances/GS.h>
#include <goofit/PDFs/physics/resonances/NonRes.h>
#include <goofit/PDFs/physics/resonances/RBW.h>

#include <cmath>

namespace GooFit {

// inDalitz, src/PDFs/physics/DalitzPlotHelpers.cu:14-59
bool inDalitz(const fptype &m12, const fptype &m13, const fptype &bigM, const fptype &dm1, const fptype &dm2,
              const fptype &dm3) {
    fptype dm1pdm2  = dm1 + dm2;
    fptype bigMmdm3 = bigM - dm3;
    fptype bigMmdm1 = bigM - dm1;
    fptype m23      = bigM * bigM + dm1 * dm1 + dm2 * dm2 + dm3 * dm3 - m12 - m13;
    if(m23 < 0. || m23 > bigMmdm1 * bigMmdm1)
        return false;
    bool m12less     = !(m12 < dm1pdm2 * dm1pdm2);
    bool m12grea     = !(m12 > bigMmdm3 * bigMmdm3);
    fptype sqrtM12   = std::sqrt(m12);
    fptype e1star    = 0.5 * (m12 - dm2 * dm2 + dm1 * dm1) / sqrtM12;
    fptype e3star    = 0.5 * (bigM * bigM - m12 - dm3 * dm3) / sqrtM12;
    fptype rte1mdm11 = std::sqrt(e1star * e1star - dm1 * dm1);
    fptype rte3mdm33 = std::sqrt(e3star * e3star - dm3 * dm3);
    fptype minimum   = (e1star + e3star) * (e1star + e3star) - (rte1mdm11 + rte3mdm33) * (rte1mdm11 + rte3mdm33);
    fptype maximum   = (e1star + e3star) * (e1star + e3star) - (rte1mdm11 - rte3mdm33) * (rte1mdm11 - rte3mdm33);
    return m12less && m12grea && !(m13 < minimum) && !(m13 > maximum);
}

// ---------------------------------------------------------------- resonances

namespace Resonances {

RBW::RBW(std::string name, Variable ar, Variable ai, Variable mass, Variable width, unsigned int sp, unsigned int cyc,
         bool norm, bool sym)
    : ResonancePdf("RBW", name, ar, ai) {
    registerParameter(mass);
    registerParameter(width);
    registerConstant(sp);
    registerConstant(cyc);
    registerConstant(norm);
    registerFunction(sym ? "ptr_to_RBW_Sym" : "ptr_to_RBW", sym ? GF_FN_RBW_SYM : GF_FN_RBW);
}

GS::GS(std::string name, Variable ar, Variable ai, Variable mass, Variable width, unsigned int sp, unsigned int cyc,
       bool norm, bool)
    : ResonancePdf("GS", name, ar, ai) {
    registerParameter(mass);
    registerParameter(width);
    registerConstant(sp);
    registerConstant(cyc);
    registerConstant(norm);
    registerFunction("ptr_to_GOUSAK", GF_FN_GS);
}

NonRes::NonRes(std::string name, Variable ar, Variable ai)
    : ResonancePdf("NonRes", name, ar, ai) {
    registerFunction("ptr_to_NONRES", GF_FN_NONRES);
}

}  // namespace Resonances

// ---------------------------------------------------------------- Amp3Body

int Amp3Body::cacheCount = 0;

Amp3Body::Amp3Body(std::string n, Observable m12, Observable m13, EventNumber eventNumber, DecayInfo3 decay,
                   GooPdf *efficiency)
    : Amp3BodyBase("Amp3Body", n, m12, m13, eventNumber)
    , decayInfo(decay)
    , _m12(m12)
    , _m13(m13)
    , _eventNumber(eventNumber) {
    if(decay.resonances.size() > 16)
        throw GeneralError("Amp3Body {}: {} resonances (at most 16)", n, decay.resonances.size());
    if(efficiency == nullptr)
        throw GeneralError("Amp3Body {}: an efficiency PDF is required", n);
    const double d6[6] = {decay.motherMass, decay.daug1Mass,    decay.daug2Mass,
                          decay.daug3Mass,  decay.meson_radius, decay.mother_meson_radius};
    engine::set_decay(d6);
    registerConstant((fptype)decayInfo.resonances.size());
    cacheToUse = cacheCount++;
    registerConstant(cacheToUse);
    for(auto *res : decayInfo.resonances) {
        registerParameter(res->amp_real);
        registerParameter(res->amp_imag);
        components.push_back(res);
    }
    components.push_back(efficiency);
    registerFunction("ptr_to_DalitzPlot", GF_FN_DALITZ);
    initialize();
    redoIntegral.assign(decayInfo.resonances.size(), true);
    integrals.assign(decayInfo.resonances.size() * decayInfo.resonances.size(), {0, 0});
    setSeparateNorm();
}

// The efficiency is the last function of this subtree (Amp3Body.cu:198-202).
void Amp3Body::populateArrays() {
    PdfBase::populateArrays();
    efficiencyFunction = (unsigned int)host_function_table.size() - 1;
}

// Amp3Body::setDataSize, src/PDFs/physics/Amp3Body.cu:204-236
void Amp3Body::setDataSize(unsigned int dataSize, unsigned int evtSize, unsigned int offset) {
    totalEventSize = (int)evtSize;
    if(totalEventSize < 3)
        throw GeneralError("totalEventSize {} must be 3 or more", totalEventSize);
    numEntries  = dataSize;
    eventOffset = (int)offset;
    engine::flush();
    engine::check(gf_dalitz_set_data_size(engine::ctx(), (int)functionIdx, dataSize, offset),
                  "gf_dalitz_set_data_size");
    cacheReady = true;
    setForceIntegrals();
}

// Amp3Body::normalize, src/PDFs/physics/Amp3Body.cu:238-359: recompute the waves of
// resonances whose mass/width changed, the integrals involving them, and the norm
// Re sum_ij c_i conj(c_j) I_ij * binSize12 * binSize13.
fptype Amp3Body::normalize() {
    recursiveSetNormalization(1.0);
    const size_t nres = decayInfo.resonances.size();
    uint64_t mask     = 0;
    bool moved        = false;  // a lineshape parameter changed since the last integrals
    for(size_t i = 0; i < nres; ++i) {
        const bool changed = decayInfo.resonances[i]->parametersChanged();
        moved              = moved || changed;
        redoIntegral[i]    = forceRedoIntegrals || changed;
        if(redoIntegral[i])
            mask |= (uint64_t)1 << i;
    }
    forceRedoIntegrals = false;
    const bool keep_integrals = cacheOnly && !moved && haveIntegrals;
    cacheOnly                 = false;
    if(mask) {
        engine::flush();
        if(cacheReady) {
            engine::check(gf_dalitz_cache(engine::ctx(), (int)functionIdx, mask), "gf_dalitz_cache");
            engine::bump_cache_epoch();
        }
    }
    if(mask && !keep_integrals) {
        std::vector<double> out(2 * nres * nres);
        if(normPoints > 0) {
            engine::check(gf_dalitz_gram_norm_points(engine::ctx(), (int)functionIdx, (int)efficiencyFunction, 0,
                                                     normPoints, out.data()),
                          "gf_dalitz_gram_norm_points");
        } else {
            const double grid6[6] = {_m12.getLowerLimit(), _m12.getUpperLimit(), (double)_m12.getNumBins(),
                                     _m13.getLowerLimit(), _m13.getUpperLimit(), (double)_m13.getNumBins()};
            engine::check(gf_dalitz_gram(engine::ctx(), (int)functionIdx, grid6, (int)efficiencyFunction, out.data()),
                          "gf_dalitz_gram");
        }
        for(size_t k = 0; k < nres * nres; k++)
            integrals[k] = {out[2 * k], out[2 * k + 1]};
        haveIntegrals = true;
    }
    // sum_ij c_i conj(c_j) I_ij in the reference's order, each complex product
    // written out as (a + ib)(c + id) = (ac - bd) + i(ad + bc), which is what the
    // std::complex products compute for finite values (this runs once per FCN point)
    if(parametersIdx + 2 * nres >= host_parameters.size())
        throw GeneralError("Amp3Body::normalize: parameter block outside the parameter array");
    const fptype *cp = host_parameters.data() + parametersIdx + 1;
    const auto *I    = reinterpret_cast<const fptype *>(integrals.data());
    fptype sum_re = 0;  // the imaginary part is not used
    for(size_t i = 0; i < nres; ++i) {
        const fptype ar = cp[2 * i], ai = cp[2 * i + 1];
        for(size_t j = 0; j < nres; ++j) {
            const fptype br = cp[2 * j], bi = -cp[2 * j + 1];
            const fptype pr = ar * br - ai * bi, pi = ar * bi + ai * br;
            const fptype ir = I[2 * (i * nres + j)], ii = I[2 * (i * nres + j) + 1];
            sum_re += pr * ir - pi * ii;
        }
    }
    fptype ret           = sum_re;
    double binSizeFactor = 1;
    if(normPoints == 0) {  // Monte-Carlo weights already carry the measure
        binSizeFactor *= _m12.getBinSize();
        binSizeFactor *= _m13.getBinSize();
    }
    ret *= binSizeFactor;
    setNormalization(1.0 / ret);
    return ret;
}

// normalize()'s last step as a device op: 1 / (Re sum_ij c_i conj(c_j) I_ij * binSizeFactor)
// over the integrals of the last normalize(); the subtree below keeps normalisation 1.
bool Amp3Body::batchNormOps(std::vector<gf_norm_op> &ops, std::vector<double> &table) const {
    if(fitControl && !fitControl->metricIsPdf())
        return true;
    const size_t nres = decayInfo.resonances.size();
    if(nres == 0 || integrals.size() != nres * nres)
        return false;
    double binSizeFactor = 1;
    if(normPoints == 0) {
        binSizeFactor *= _m12.getBinSize();
        binSizeFactor *= _m13.getBinSize();
    }
    gf_norm_op op{};
    op.kind = GF_NORM_DALITZ;
    op.out  = (int32_t)normalIdx + 1;
    op.p    = (int32_t)parametersIdx + 1;
    op.n    = (int32_t)nres;
    op.g    = (int32_t)table.size();
    op.c0   = binSizeFactor;
    for(const auto &v : integrals) {
        table.push_back(v.real());
        table.push_back(v.imag());
    }
    ops.push_back(op);
    return true;
}

void Amp3Body::setNormalisationPoints(int64_t n, unsigned long long seed, DalitzProposal *proposal) {
    if(n < 1)
        throw GeneralError("Amp3Body::setNormalisationPoints: need at least one point");
    if(proposal && !proposal->isBuilt())
        throw GeneralError("Amp3Body::setNormalisationPoints: the proposal has no table (DalitzProposal::build)");
    const double box[4] = {_m12.getLowerLimit(), _m12.getUpperLimit(), _m13.getLowerLimit(), _m13.getUpperLimit()};
    auto draw = [&]() {
        engine::flush();
        engine::check(gf_dalitz_set_norm_points(engine::ctx(), (int)functionIdx,
                                                proposal ? proposal->handle() : nullptr, n, seed, box),
                      "gf_dalitz_set_norm_points");
        normPoints         = n;
        forceRedoIntegrals = true;
        cacheOnly          = false;
    };
    draw();
    if(!proposal || !proposal->retrainingEnabled())
        return;
    // the retraining trigger: the weight variance of these points at the current couplings
    // against the variance the proposal had when it was (re)built for the model
    copyParams();
    const int batches = n >= 16 ? 16 : (int)n;
    if(batches < 2)
        return;
    const fptype relvar = normalisationRelativeVariance(batches);
    if(proposal->getReferenceVariance() <= 0) {
        proposal->setReferenceVariance(relvar);
        return;
    }
    if(!proposal->needsRetraining(relvar, proposal->retraining().factor))
        return;
    // train again on events of the model at its current parameters, then redraw
    proposal->retrain(*this, seed);
    draw();
    proposal->setReferenceVariance(normalisationRelativeVariance(batches));
}

void Amp3Body::setNormalisationRule(const std::vector<fptype> &m12, const std::vector<fptype> &m13,
                                    const std::vector<fptype> &w) {
    if(m12.empty() || m12.size() != m13.size() || m12.size() != w.size())
        throw GeneralError("Amp3Body::setNormalisationRule: need equal, non-empty m12, m13 and weight arrays");
    engine::flush();
    engine::check(gf_dalitz_set_norm_rule(engine::ctx(), (int)functionIdx, m12.data(), m13.data(), w.data(),
                                          (int64_t)m12.size()),
                  "gf_dalitz_set_norm_rule");
    normPoints         = (int64_t)m12.size();
    forceRedoIntegrals = true;
    cacheOnly          = false;
}

void Amp3Body::setNormalisationGrid() {
    normPoints         = 0;
    forceRedoIntegrals = true;
    cacheOnly          = false;
}

fptype Amp3Body::normalisationRelativeVariance(int batches) {
    if(normPoints == 0)
        throw GeneralError("Amp3Body::normalisationRelativeVariance: normalisation is on the grid");
    if(batches < 2 || normPoints / batches < 1)
        throw GeneralError("Amp3Body::normalisationRelativeVariance: need at least 2 non-empty batches");
    engine::flush();
    const size_t nres   = decayInfo.resonances.size();
    const int64_t count = normPoints / batches;
    std::vector<double> out(2 * nres * nres);
    std::vector<double> est(batches);
    for(int b = 0; b < batches; b++) {
        engine::check(gf_dalitz_gram_norm_points(engine::ctx(), (int)functionIdx, (int)efficiencyFunction,
                                                 (int64_t)b * count, count, out.data()),
                      "gf_dalitz_gram_norm_points");
        std::complex<fptype> sum(0, 0);
        for(size_t i = 0; i < nres; ++i) {
            std::complex<fptype> ai(host_parameters.at(parametersIdx + i * 2 + 1),
                                    host_parameters.at(parametersIdx + i * 2 + 2));
            for(size_t j = 0; j < nres; ++j) {
                std::complex<fptype> aj(host_parameters.at(parametersIdx + j * 2 + 1),
                                        -host_parameters.at(parametersIdx + j * 2 + 2));
                sum += ai * aj * std::complex<fptype>(out[2 * (i * nres + j)], out[2 * (i * nres + j) + 1]);
            }
        }
        est[b] = sum.real() * (double)normPoints / (double)count;  // the batch's own estimate of the integral
    }
    double mean = 0, var = 0;
    for(double v : est)
        mean += v;
    mean /= batches;
    for(double v : est)
        var += (v - mean) * (v - mean);
    var /= (batches - 1);
    return var / (mean * mean) * (double)count;
}

int64_t Amp3Body::generateEvents(UnbinnedDataSet &out, size_t n, unsigned long long seed, double fmax) {
    if(n == 0)
        return 0;
    const double box[4] = {_m12.getLowerLimit(), _m12.getUpperLimit(), _m13.getLowerLimit(), _m13.getUpperLimit()};
    std::vector<fptype> a(n), b(n), num(n);
    engine::flush();
    int64_t tried = 0;
    engine::check(gf_dalitz_generate(engine::ctx(), (int)functionIdx, (int64_t)n, seed, box, fmax, 1, a.data(),
                                     b.data(), &tried, nullptr),
                  "gf_dalitz_generate");
    const size_t first = out.getNumEvents();
    for(size_t i = 0; i < n; i++)
        num[i] = (fptype)(first + i);
    std::vector<const fptype *> cols;
    for(const Observable &o : out.getObservables()) {
        if(o == _m12)
            cols.push_back(a.data());
        else if(o == _m13)
            cols.push_back(b.data());
        else if(o == _eventNumber)
            cols.push_back(num.data());
        else
            throw GeneralError("Amp3Body::generateEvents: the data set has observable {} besides m12, m13 and the "
                               "event number",
                               o.getName());
    }
    if(cols.size() != 3)
        throw GeneralError("Amp3Body::generateEvents: the data set needs m12, m13 and the event number");
    out.appendColumns(cols, n);
    return tried;
}

fptype Amp3Body::integralAtPoints(const std::vector<fptype> &m12, const std::vector<fptype> &m13,
                                  const std::vector<fptype> &w) {
    if(m12.size() != m13.size() || m12.size() != w.size() || m12.empty())
        throw GeneralError("Amp3Body::integralAtPoints: need equal, non-empty m12, m13 and weight arrays");
    copyParams();
    engine::flush();
    const size_t nres = decayInfo.resonances.size();
    std::vector<double> out(2 * nres * nres);
    engine::check(gf_dalitz_gram_points(engine::ctx(), (int)functionIdx, m12.data(), m13.data(), w.data(),
                                        (int64_t)m12.size(), (int)efficiencyFunction, out.data()),
                  "gf_dalitz_gram_points");
    const fptype *cp = host_parameters.data() + parametersIdx + 1;
    fptype sum_re    = 0;
    for(size_t i = 0; i < nres; ++i) {
        const fptype ar = cp[2 * i], ai = cp[2 * i + 1];
        for(size_t j = 0; j < nres; ++j) {
            const fptype br = cp[2 * j], bi = -cp[2 * j + 1];
            const fptype pr = ar * br - ai * bi, pi = ar * bi + ai * br;
            sum_re += pr * out[2 * (i * nres + j)] - pi * out[2 * (i * nres + j) + 1];
        }
    }
    return sum_re;
}

const std::vector<std::complex<fptype>> Amp3Body::getCachedWave(size_t i) const {
    engine::flush();
    std::vector<std::complex<fptype>> out(numEntries);
    engine::check(gf_dalitz_get_cache(engine::ctx(), (int)functionIdx, (int)i, reinterpret_cast<double *>(out.data()),
                                      numEntries),
                  "gf_dalitz_get_cache");
    return out;
}

std::complex<fptype> Amp3Body::sumCachedWave(size_t i) const {
    auto w = getCachedWave(i);
    std::complex<fptype> s(0, 0);
    for(auto &v : w)
        s += v;
    return s;
}

// Amp3Body::fit_fractions, src/PDFs/physics/Amp3Body.cu:376-428, with the n x n
// sums over the cache done as one Gram G_ij = sum_e W_i conj(W_j).
std::vector<std::vector<fptype>> Amp3Body::fit_fractions() {
    const size_t n = decayInfo.resonances.size();
    engine::flush();
    std::vector<double> G(2 * n * n);
    engine::check(gf_dalitz_cache_gram(engine::ctx(), (int)functionIdx, G.data()), "gf_dalitz_cache_gram");
    std::vector<std::complex<fptype>> c(n);
    for(size_t i = 0; i < n; i++)
        c[i] = {decayInfo.resonances[i]->amp_real.getValue(), decayInfo.resonances[i]->amp_imag.getValue()};
    std::vector<std::vector<fptype>> amps(n, std::vector<fptype>(n));
    fptype total = 0;
    for(size_t i = 0; i < n; i++)
        for(size_t j = 0; j < n; j++) {
            std::complex<fptype> g(G[2 * (i * n + j)], G[2 * (i * n + j) + 1]);
            fptype v   = (c[i] * std::conj(c[j]) * g).real() / numEntries;
            amps[i][j] = v;
            total += v;
        }
    std::vector<std::vector<fptype>> ff(n, std::vector<fptype>(n));
    for(size_t i = 0; i < n; i++)
        for(size_t j = 0; j < n; j++)
            ff[i][j] = amps[i][j] / total;
    return ff;
}

}  // namespace GooFit
