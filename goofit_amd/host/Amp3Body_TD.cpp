// Amp3Body_TD.cpp -- the time-dependent three-body Dalitz PDF and its decay-time
// resolutions (src/PDFs/physics/Amp3Body_TD.cu, TruthResolution.cu,
// ThreeGaussResolution.cu and MixingTimeResolution.cu of the reference): tree
// layout, wave-cache management and the host assembly of the normalisation.
#include <goofit/Error.h>
#include <goofit/PDFs/physics/Amp3Body_TD.h>
#include <goofit/PDFs/physics/ThreeGaussResolution.h>
#include <goofit/PDFs/physics/TruthResolution.h>

#include <cstdio>

namespace GooFit {

// ---------------------------------------------------------------- resolutions

MixingTimeResolution::~MixingTimeResolution() = default;

TruthResolution::TruthResolution()
    : MixingTimeResolution("TruthResolution") {
    registerFunction("ptr_to_truth", GF_FN_TRUTH_RES);
}

TruthResolution::~TruthResolution() = default;

// TruthResolution::normalization, src/PDFs/physics/TruthResolution.cu:58-71: the
// decay-time integrals of e^{-t/tau} times cosh, cos, sinh, sin.
fptype TruthResolution::normalization(fptype di1, fptype di2, fptype di3, fptype di4, fptype tau, fptype xmixing,
                                      fptype ymixing) const {
    fptype timeIntegralOne = tau / (1 - ymixing * ymixing);
    fptype timeIntegralTwo = tau / (1 + xmixing * xmixing);
    fptype timeIntegralThr = ymixing * timeIntegralOne;
    fptype timeIntegralFou = xmixing * timeIntegralTwo;
    fptype ret             = timeIntegralOne * (di1 + di2);
    ret += timeIntegralTwo * (di1 - di2);
    ret -= 2 * timeIntegralThr * di3;
    ret -= 2 * timeIntegralFou * di4;
    return ret;
}

ThreeGaussResolution::ThreeGaussResolution(Variable cf, Variable tf, Variable cb, Variable cs, Variable tb, Variable ts,
                                           Variable ob, Variable os, Variable sb, Variable common_bias_offset,
                                           Variable common_width_scale, Variable common_bias_offset2,
                                           Variable common_width_scale2)
    : MixingTimeResolution("ThreeGaussResolution", cf, tf, cb, cs, tb, ts, ob, os, sb)
    , selectionBias(sb) {
    registerParameter(common_bias_offset);
    registerParameter(common_width_scale);
    registerParameter(common_bias_offset2);
    registerParameter(common_width_scale2);
    registerFunction("ptr_to_threegauss", GF_FN_THREEGAUSS_RES);
}

ThreeGaussResolution::~ThreeGaussResolution() = default;

// ThreeGaussResolution::normalization, src/PDFs/physics/ThreeGaussResolution.cu:155-182:
// the decay-time integrals with the selection bias s, e^{-(1/tau + s) t}.
fptype ThreeGaussResolution::normalization(fptype di1, fptype di2, fptype di3, fptype di4, fptype tau, fptype xmixing,
                                           fptype ymixing) const {
    fptype selBias = selectionBias.getValue();
    fptype timeIntegralOne
        = (selBias + 1 / tau) / (selBias * selBias + 2 * selBias / tau + (1 - ymixing * ymixing) / (tau * tau));
    fptype timeIntegralTwo
        = (selBias + 1 / tau) / (selBias * selBias + 2 * selBias / tau + (1 + xmixing * xmixing) / (tau * tau));
    fptype timeIntegralThr
        = (ymixing / tau) / (selBias * selBias + 2 * selBias / tau + (1 - ymixing * ymixing) / (tau * tau));
    fptype timeIntegralFou
        = (xmixing / tau) / (selBias * selBias + 2 * selBias / tau + (1 + xmixing * xmixing) / (tau * tau));
    fptype ret = timeIntegralOne * (di1 + di2);
    ret += timeIntegralTwo * (di1 - di2);
    ret -= 2 * timeIntegralThr * di3;
    ret -= 2 * timeIntegralFou * di4;
    return ret;
}

// ---------------------------------------------------------------- Amp3Body_TD

int Amp3Body_TD::cacheCount = 0;

// Amp3Body_TD::Amp3Body_TD, src/PDFs/physics/Amp3Body_TD.cu:262-359: parameters
// tau, x, y, dx, dy and the couplings; constants nres, cacheToUse, the mistag
// flag and the function index after the efficiency; components the
// resonances, the resolution and the efficiency.
Amp3Body_TD::Amp3Body_TD(std::string n, Observable _dtime, Observable _sigmat, Observable m12, Observable m13,
                         EventNumber eventNumber, DecayInfo3t decay, MixingTimeResolution *r, GooPdf *efficiency,
                         Observable *mistag, Observable *charmtag)
    : Amp3BodyBase("Amp3Body_TD", n, _dtime, _sigmat, m12, m13, eventNumber)
    , decayInfo(decay)
    , _m12(m12)
    , _m13(m13)
    , resolution(r)
    , hasMistag(mistag != nullptr)
    , _mistag(mistag ? *mistag : Observable("no_mistag", 0, 0)) {
    if(decay.resonances.size() > 16)
        throw GeneralError("Amp3Body_TD {}: {} resonances (at most 16)", n, decay.resonances.size());
    if(efficiency == nullptr)
        throw GeneralError("Amp3Body_TD {}: an efficiency PDF is required", n);
    if(resolution == nullptr || resolution->getDeviceFunction() < 0)
        throw GeneralError("Amp3Body_TD {}: the resolution has no device function", n);
    if(charmtag && !mistag)
        throw GeneralError("Amp3Body_TD {}: a charm tag needs the mistag observable before it", n);
    if(mistag)
        registerObservable(*mistag);
    if(charmtag)
        registerObservable(*charmtag);
    const double d6[6] = {decay.motherMass, decay.daug1Mass,    decay.daug2Mass,
                          decay.daug3Mass,  decay.meson_radius, decay.mother_meson_radius};
    engine::set_decay(d6);
    registerParameter(decay._tau);
    registerParameter(decay._xmixing);
    registerParameter(decay._ymixing);
    registerParameter(decay._deltax);
    registerParameter(decay._deltay);
    setD0Fraction(0.5);
    registerConstant((fptype)decayInfo.resonances.size());
    cacheToUse = cacheCount++;
    registerConstant(cacheToUse);
    registerConstant(mistag == nullptr ? 1 : 0);
    for(auto *res : decayInfo.resonances) {
        registerParameter(res->amp_real);
        registerParameter(res->amp_imag);
        components.push_back(res);
    }
    components.push_back(resolution);
    components.push_back(efficiency);
    registerConstant(0);  // the function index after the efficiency, set by populateArrays
    registerFunction("ptr_to_Tddp", GF_FN_TDDP);
    initialize();
    const size_t nres = decayInfo.resonances.size();
    redoIntegral.assign(nres, true);
    iaa.assign(nres * nres, {0, 0});
    iab.assign(nres * nres, {0, 0});
    ibb.assign(nres * nres, {0, 0});
    setSeparateNorm();
}

Amp3Body_TD::Amp3Body_TD(std::string n, Observable, Observable, Observable m12, Observable m13, EventNumber,
                         DecayInfo3t decay, std::vector<MixingTimeResolution *> &, GooPdf *, Observable,
                         Observable *, Observable *)
    : Amp3BodyBase("Amp3Body_TD", n)
    , decayInfo(decay)
    , _m12(m12)
    , _m13(m13)
    , _mistag("no_mistag", 0, 0) {
    throw GeneralError("Amp3Body_TD {}: per-D0-mass-bin resolutions are not supported (the reference's device "
                       "function never selects among them)",
                       n);
}

// Amp3Body_TD::populateArrays, src/PDFs/physics/Amp3Body_TD.cu:472-529: the
// resolution follows the resonances, the efficiency subtree follows it, and the
// last constant becomes the function index after the efficiency.
void Amp3Body_TD::populateArrays() {
    PdfBase::populateArrays();
    const size_t nres  = decayInfo.resonances.size();
    resolutionFunction = functionIdx + 1 + (unsigned int)nres;
    efficiencyFunction = resolutionFunction + 1;
    constantsList.back()                                    = (fptype)host_function_table.size();
    host_constants.at(constantsIdx + constantsList.size()) = constantsList.back();
}

// Amp3Body_TD::setDataSize, src/PDFs/physics/Amp3Body_TD.cu:531-580
void Amp3Body_TD::setDataSize(unsigned int dataSize, unsigned int evtSize, unsigned int offset) {
    totalEventSize = (int)evtSize;
    if(totalEventSize < 5)
        throw GeneralError("totalEventSize {} must be 5 or more", totalEventSize);
    numEntries  = dataSize;
    eventOffset = (int)offset;
    engine::flush();
    engine::check(gf_dalitz_set_data_size(engine::ctx(), (int)functionIdx, dataSize, offset),
                  "gf_dalitz_set_data_size");
    cacheReady = true;
    setForceIntegrals();
}

void Amp3Body_TD::setD0Fraction(fptype d0fraction) {
    if(!(d0fraction >= 0 && d0fraction <= 1))
        throw GeneralError("Amp3Body_TD: D0 fraction {} outside [0, 1]", d0fraction);
    _D0Fraction = d0fraction;
}

fptype Amp3Body_TD::getD0Fraction() { return _D0Fraction; }

// Waves of the resonances in mask (when the cache is allocated), then every
// ThreeComplex integral over the grid (one stacked Gram).
void Amp3Body_TD::computeIntegrals(uint64_t mask) {
    const size_t nres = decayInfo.resonances.size();
    engine::flush();
    if(cacheReady && mask) {
        engine::check(gf_dalitz_cache(engine::ctx(), (int)functionIdx, mask), "gf_dalitz_cache");
        engine::bump_cache_epoch();
    }
    const double grid6[6] = {_m12.getLowerLimit(), _m12.getUpperLimit(), (double)_m12.getNumBins(),
                             _m13.getLowerLimit(), _m13.getUpperLimit(), (double)_m13.getNumBins()};
    std::vector<double> a(2 * nres * nres), b(2 * nres * nres), c(2 * nres * nres);
    engine::check(gf_tddp_gram(engine::ctx(), (int)functionIdx, grid6, a.data(), b.data(), c.data()),
                  "gf_tddp_gram");
    for(size_t k = 0; k < nres * nres; k++) {
        iaa[k] = {a[2 * k], a[2 * k + 1]};
        iab[k] = {b[2 * k], b[2 * k + 1]};
        ibb[k] = {c[2 * k], c[2 * k + 1]};
    }
}

// Amp3Body_TD::normalize, src/PDFs/physics/Amp3Body_TD.cu:590-751: the Dalitz
// integrals |A|^2, |B|^2, A B* from the ThreeComplex integrals and the couplings,
// the resolution's decay-time integral for D0 (x + dx, y + dy) and D0bar
// (x - dx, y - dy) mixed by the D0 fraction, the mistag terms, and the bin area.
fptype Amp3Body_TD::normalize() {
    recursiveSetNormalization(1.0);
    const size_t nres = decayInfo.resonances.size();
    uint64_t mask     = 0;
    for(size_t i = 0; i < nres; ++i) {
        redoIntegral[i] = forceRedoIntegrals || decayInfo.resonances[i]->parametersChanged();
        if(redoIntegral[i])
            mask |= (uint64_t)1 << i;
    }
    forceRedoIntegrals = false;
    if(mask)
        computeIntegrals(mask);
    std::complex<fptype> integralA_2(0, 0), integralB_2(0, 0), integralABs(0, 0);
    for(size_t i = 0; i < nres; ++i) {
        std::complex<fptype> amplitude_i(host_parameters.at(parametersIdx + i * 2 + 6),
                                         host_parameters.at(parametersIdx + i * 2 + 7));
        for(size_t j = 0; j < nres; ++j) {
            std::complex<fptype> amplitude_j(host_parameters.at(parametersIdx + j * 2 + 6),
                                             -host_parameters.at(parametersIdx + j * 2 + 7));
            integralA_2 += amplitude_i * amplitude_j * iaa[i * nres + j];
            integralABs += amplitude_i * amplitude_j * iab[i * nres + j];
            integralB_2 += amplitude_i * amplitude_j * ibb[i * nres + j];
        }
    }
    const double dalitzIntegralOne = integralA_2.real();
    const double dalitzIntegralTwo = integralB_2.real();
    const double dalitzIntegralThr = integralABs.real();
    const double dalitzIntegralFou = integralABs.imag();
    const fptype tau      = host_parameters.at(parametersIdx + 1);
    const fptype xmixing0 = host_parameters.at(parametersIdx + 2);
    const fptype ymixing0 = host_parameters.at(parametersIdx + 3);
    const fptype deltax   = host_parameters.at(parametersIdx + 4);
    const fptype deltay   = host_parameters.at(parametersIdx + 5);
    fptype ret_D0    = resolution->normalization(dalitzIntegralOne, dalitzIntegralTwo, dalitzIntegralThr,
                                              dalitzIntegralFou, tau, xmixing0 + deltax, ymixing0 + deltay);
    fptype ret_D0bar = resolution->normalization(dalitzIntegralOne, dalitzIntegralTwo, dalitzIntegralThr,
                                                 dalitzIntegralFou, tau, xmixing0 - deltax, ymixing0 - deltay);
    fptype ret       = _D0Fraction * ret_D0 + (1. - _D0Fraction) * ret_D0bar;
    const fptype xfix = 0.0039, yfix = 0.0065, taufix = 0.4101;
    // the mistag observable tests true when its host value is inside its limits
    if(hasMistag && static_cast<bool>(_mistag)) {
        const fptype w = _mistag.getValue();
        ret *= (1 - 2 * w);
        ret += w * resolution->normalization(dalitzIntegralOne, dalitzIntegralTwo, dalitzIntegralThr,
                                             dalitzIntegralFou, taufix, xfix, yfix);
        ret += w * resolution->normalization(dalitzIntegralTwo, dalitzIntegralOne, dalitzIntegralThr,
                                             -dalitzIntegralFou, taufix, xfix, yfix);
    }
    double binSizeFactor = 1;
    binSizeFactor *= _m12.getBinSize();
    binSizeFactor *= _m13.getBinSize();
    ret *= binSizeFactor;
    setNormalization(1.0 / ret);
    return ret;
}

std::vector<WaveHolder_s> Amp3Body_TD::getCachedWaveNoCopy(size_t i) const {
    const size_t nres = decayInfo.resonances.size();
    if(i >= nres)
        throw GeneralError("Amp3Body_TD::getCachedWaveNoCopy: resonance {} of {}", i, nres);
    engine::flush();
    std::vector<std::complex<fptype>> a(numEntries), b(numEntries);
    engine::check(gf_dalitz_get_cache(engine::ctx(), (int)functionIdx, (int)i, reinterpret_cast<double *>(a.data()),
                                      numEntries),
                  "gf_dalitz_get_cache");
    engine::check(gf_dalitz_get_cache(engine::ctx(), (int)functionIdx, (int)(nres + i),
                                      reinterpret_cast<double *>(b.data()), numEntries),
                  "gf_dalitz_get_cache");
    std::vector<WaveHolder_s> out(numEntries);
    for(size_t e = 0; e < numEntries; e++)
        out[e] = {a[e].real(), a[e].imag(), b[e].real(), b[e].imag()};
    return out;
}

// Amp3Body_TD::getFractions, src/PDFs/physics/Amp3Body_TD.cu:754-910
std::vector<std::vector<fptype>> Amp3Body_TD::getFractions() {
    recursiveSetNormalization(1.0);
    const size_t nres = decayInfo.resonances.size();
    uint64_t mask     = 0;
    for(size_t i = 0; i < nres; ++i)
        if(forceRedoIntegrals || decayInfo.resonances[i]->parametersChanged())
            mask |= (uint64_t)1 << i;
    forceRedoIntegrals = false;
    if(mask)
        computeIntegrals(mask);
    std::complex<fptype> integralA_2(0, 0);
    std::vector<fptype> matdiag(nres);
    std::vector<std::vector<fptype>> matdiag_int(nres, std::vector<fptype>(nres));
    for(size_t i = 0; i < nres; ++i) {
        std::complex<fptype> amplitude_i(host_parameters.at(parametersIdx + i * 2 + 6),
                                         host_parameters.at(parametersIdx + i * 2 + 7));
        for(size_t j = 0; j < nres; ++j) {
            std::complex<fptype> amplitude_j(host_parameters.at(parametersIdx + j * 2 + 6),
                                             -host_parameters.at(parametersIdx + j * 2 + 7));
            const std::complex<fptype> term = amplitude_i * amplitude_j * iaa[i * nres + j];
            integralA_2 += term;
            if(i == j)
                matdiag[i] = term.real();
            matdiag_int[i][j] = term.real();
        }
    }
    for(size_t i = 0; i < nres; ++i)
        printf("Integral contribution for res # %zu ( %s) : %.4g%%.\n", i, decayInfo.resonances[i]->getName().c_str(),
               matdiag[i] / integralA_2.real() * 100.);
    for(size_t i = 0; i < nres; ++i)
        for(size_t j = i + 1; j < nres; ++j)
            printf("Integral interference fraction for resonances %zu, %zu ( %s, %s) : %.4g%%.\n", i, j,
                   decayInfo.resonances[i]->getName().c_str(), decayInfo.resonances[j]->getName().c_str(),
                   matdiag_int[i][j] / integralA_2.real() * 100.);
    return matdiag_int;
}

}  // namespace GooFit
