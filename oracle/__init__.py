"""CPU oracle for GooFit's likelihood hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this package, and only as the checker.  Nothing in the
product (``goofit_amd``) imports it.

Two parts:

* ``liboracle.so`` (``goofit_oracle.c``): the device-side algorithm restated in
  C -- the flattened-tree walk, every PDF function, the metrics, OpenMP-chunked
  reductions, the 1-D grid integral and the Dalitz lineshapes and integrals.
* This module: a ctypes wrapper plus a Python restatement of the reference's
  host-side logic -- building the flattened arrays
  (``PdfBase::populateArrays``, src/PDFs/PdfBase.cu:268-311; observable
  numbering ``PdfBase::setupObservables``, src/goofit/PdfBase.cpp:109-117),
  ``normalize()`` of GooPdf / AddPdf / ProdPdf / Amp3Body
  (src/PDFs/GooPdf.cu:340-396, src/PDFs/combine/AddPdf.cu:143-185,
  src/PDFs/combine/ProdPdf.cu:75-99, src/PDFs/physics/Amp3Body.cu:238-359)
  and ``calculateNLL()`` (src/PDFs/GooPdf.cu:254-273, AddPdf.cu:187-204).
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "liboracle.so")

# Function codes: the gf_func enum of include/goofit_hip.h.
FN = dict(NONE=0, GAUSSIAN=1, EXP=2, EXP_OFFSET=3, EXP_POLY=4, EXP_POLY_OFFSET=5, EXPGAUS=6,
          POLYNOMIAL=7, OFFSET_POLYNOMIAL=8, MULTI_POLYNOMIAL=9, ADD=10, ADD_EXT=11, PROD=12,
          DALITZ=13, RBW=14, RBW_SYM=15, GS=16, NONRES=17, TDDP=18, TRUTH_RES=19, THREEGAUSS_RES=20)
METRIC = dict(EVAL=0, NLL=1, PROB=2, BINAVG=3, BINWITHERROR=4, CHISQ=5)

_lib = None

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int32)


def build():
    """Compile liboracle.so with the recipe in oracle/Makefile."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.gfo_prog_new.restype = ctypes.c_void_p
        L.gfo_prog_new.argtypes = [_I, ctypes.c_int, _D, ctypes.c_int, _D, ctypes.c_int, _D, ctypes.c_int,
                                   _D, ctypes.c_int, _D]
        L.gfo_prog_free.argtypes = [ctypes.c_void_p]
        L.gfo_prog_set_params.argtypes = [ctypes.c_void_p, _D, ctypes.c_int]
        L.gfo_prog_set_norms.argtypes = [ctypes.c_void_p, _D, ctypes.c_int]
        L.gfo_prog_set_decay.argtypes = [ctypes.c_void_p, _D]
        L.gfo_prog_set_cache.argtypes = [ctypes.c_void_p, ctypes.c_int, _D, ctypes.c_int64]
        L.gfo_prog_set_total_events.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.gfo_prog_error.argtypes = [ctypes.c_void_p]
        L.gfo_prog_error.restype = ctypes.c_int
        L.gfo_eval_event.argtypes = [ctypes.c_void_p, _D]
        L.gfo_eval_event.restype = ctypes.c_double
        L.gfo_eval_event_at.argtypes = [ctypes.c_void_p, ctypes.c_int, _D]
        L.gfo_eval_event_at.restype = ctypes.c_double
        L.gfo_metric.argtypes = [ctypes.c_int, ctypes.c_double, _D, ctypes.c_double, ctypes.c_double]
        L.gfo_metric.restype = ctypes.c_double
        L.gfo_eval_metric.argtypes = [ctypes.c_void_p, _D, ctypes.c_int64, ctypes.c_int, ctypes.c_int, _D]
        L.gfo_reduce_metric.argtypes = [ctypes.c_void_p, _D, ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.gfo_reduce_metric.restype = ctypes.c_double
        L.gfo_chunked_sum.argtypes = [_D, ctypes.c_int64, ctypes.c_int]
        L.gfo_chunked_sum.restype = ctypes.c_double
        L.gfo_exact_sum.argtypes = [_D, ctypes.c_int64]
        L.gfo_exact_sum.restype = ctypes.c_double
        L.gfo_num_procs.restype = ctypes.c_int
        L.gfo_reduce_bins.argtypes = [ctypes.c_void_p, ctypes.c_int, _D, ctypes.c_int, ctypes.c_int]
        L.gfo_reduce_bins.restype = ctypes.c_double
        L.gfo_eval_bins.argtypes = [ctypes.c_void_p, ctypes.c_int, _D, ctypes.c_int, _D]
        L.gfo_grid_cells.argtypes = [_D, ctypes.c_int, _D]
        L.gfo_dalitz_grid_cells.argtypes = [_D, ctypes.c_int, _D, _D]
        L.gfo_in_dalitz.argtypes = [ctypes.c_double] * 6
        L.gfo_in_dalitz.restype = ctypes.c_int
        L.gfo_two_body_cm_mom.argtypes = [ctypes.c_double] * 3
        L.gfo_two_body_cm_mom.restype = ctypes.c_double
        L.gfo_lineshape.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_double, _D]
        L.gfo_dalitz_wave.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _D, ctypes.c_int64, ctypes.c_int,
                                      ctypes.c_int64, _D]
        L.gfo_dalitz_layout.argtypes = [ctypes.c_void_p, ctypes.c_int, _I, _I]
        L.gfo_dalitz_layout.restype = ctypes.c_int
        L.gfo_dalitz_integrals.argtypes = [ctypes.c_void_p, ctypes.c_int, _D, ctypes.c_int, ctypes.c_int, _D]
        L.gfo_dalitz_norm_from_integrals.argtypes = [ctypes.c_int, _D, _D, ctypes.c_double, ctypes.c_double]
        L.gfo_dalitz_norm_from_integrals.restype = ctypes.c_double
        L.gfo_prog_set_cache_b.argtypes = [ctypes.c_void_p, ctypes.c_int, _D]
        L.gfo_tddp_layout.argtypes = [ctypes.c_void_p, ctypes.c_int, _I, _I, _I, _I]
        L.gfo_tddp_layout.restype = ctypes.c_int
        L.gfo_tddp_wave.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _D, ctypes.c_int64, ctypes.c_int,
                                    ctypes.c_int64, _D, _D]
        L.gfo_tddp_integrals.argtypes = [ctypes.c_void_p, ctypes.c_int, _D, ctypes.c_int, _D, _D, _D]
        L.gfo_resolution_norm.argtypes = [ctypes.c_int] + [ctypes.c_double] * 8
        L.gfo_resolution_norm.restype = ctypes.c_double
        L.gfo_resolution_eval.argtypes = [ctypes.c_int, _D] + [ctypes.c_double] * 9
        L.gfo_resolution_eval.restype = ctypes.c_double
        L.gfo_tddp_norm_from_integrals.argtypes = [ctypes.c_int, _D, _D, _D, _D, _D, ctypes.c_int, ctypes.c_double,
                                                   ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                                   ctypes.c_double]
        L.gfo_tddp_norm_from_integrals.restype = ctypes.c_double
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(_D)


def num_procs() -> int:
    """omp_get_num_procs() -- the reference OMP build's chunk count."""
    return lib().gfo_num_procs()


def chunked_sum(terms, nchunks=None) -> float:
    t = np.ascontiguousarray(terms, dtype=np.float64)
    return lib().gfo_chunked_sum(_dp(t), t.size, nchunks or num_procs())


def exact_sum(terms) -> float:
    t = np.ascontiguousarray(terms, dtype=np.float64)
    return lib().gfo_exact_sum(_dp(t), t.size)


def grid_cells(ranges):
    """Cell centres of a k-dimensional grid, shape (k, nbins) (MetricTaker.cu:76-86)."""
    r = np.ascontiguousarray(ranges, np.float64)
    k = r.size // 3
    n = int(np.prod([int(math.floor(r[3 * i + 2] + 0.5)) for i in range(k)]))
    out = np.empty(k * n)
    lib().gfo_grid_cells(_dp(r), k, _dp(out))
    return out.reshape(k, n)


def dalitz_grid_cells(grid6, time_dependent=False):
    """Centres (m12, m13) of the Dalitz integration cells (SpecialResonanceIntegrator.cu:59-76,
    SpecialDalitzIntegrator.cu:75-82)."""
    g = np.ascontiguousarray(grid6, np.float64)
    n = int(math.floor(g[2] + 0.5)) * int(math.floor(g[5] + 0.5))
    m12, m13 = np.empty(n), np.empty(n)
    lib().gfo_dalitz_grid_cells(_dp(g), 1 if time_dependent else 0, _dp(m12), _dp(m13))
    return m12, m13


def in_dalitz(m12, m13, M, d1, d2, d3) -> bool:
    return bool(lib().gfo_in_dalitz(m12, m13, M, d1, d2, d3))


def two_body_cm_mom(s, d1, d2) -> float:
    return lib().gfo_two_body_cm_mom(s, d1, d2)


# ----------------------------------------------------------------------------
# Host-side model: a Python restatement of the reference PDF classes' layout.
# ----------------------------------------------------------------------------

class Obs:
    """Observable (include/goofit/Variable.h:109-142): limits, bins and a shared index."""

    def __init__(self, name, lo, hi, nbins=100, event_number=False):
        self.name, self.lo, self.hi, self.nbins = name, float(lo), float(hi), int(nbins)
        self.value = 0.0  # host value (Indexable(name, 0, lo, hi), include/goofit/Variable.h:120-121)
        self.index = -1
        self.event_number = event_number

    def bin_size(self):
        return (self.hi - self.lo) / self.nbins


class Var:
    """Variable (include/goofit/Variable.h:147-241)."""

    def __init__(self, name, value, err=0.1, lo=0.0, hi=0.0, fixed=False):
        self.name, self.value, self.err, self.lo, self.hi, self.fixed = name, float(value), err, lo, hi, fixed


class Pdf:
    func = FN["NONE"]

    def __init__(self, name, observables=(), params=(), consts=(), components=()):
        self.name = name
        self.observables = list(observables)
        self.params = list(params)
        self.consts = [float(c) for c in consts]
        self.components = list(components)
        self.cached_norm = 1.0
        self.function_idx = self.parameters_idx = self.constants_idx = 0
        self.observables_idx = self.normal_idx = 0
        self.integration_bins = -1

    def all_observables(self):
        out = list(self.observables)
        for c in self.components:
            for o in c.all_observables():
                if o not in out:
                    out.append(o)
        return out

    def has_analytic_integral(self):
        return False

    def integrate(self, lo, hi):
        return 0.0


class Program:
    """The four flattened arrays + function codes (PdfBase::populateArrays)."""

    def __init__(self):
        self.funcs, self.params, self.consts, self.obs, self.norms = [], [], [], [], []
        self.nodes = []

    def populate(self, pdf: Pdf):
        pdf.function_idx = len(self.funcs)
        self.funcs.append(pdf.func)
        self.nodes.append(pdf)
        pdf.parameters_idx = len(self.params)
        self.params.append(float(len(pdf.params)))
        self.params += [p.value for p in pdf.params]
        pdf.constants_idx = len(self.consts)
        self.consts.append(float(len(pdf.consts)))
        self.consts += pdf.consts
        pdf.observables_idx = len(self.obs)
        self.obs.append(float(len(pdf.observables)))
        self.obs += [float(o.index) for o in pdf.observables]
        pdf.normal_idx = len(self.norms)
        self.norms += [1.0, pdf.cached_norm]
        for c in pdf.components:
            self.populate(c)
        if pdf.func == FN["TDDP"]:
            # Amp3Body_TD::populateArrays (Amp3Body_TD.cu:523-527): the last constant is
            # the function index after the efficiency subtree
            pdf.consts[-1] = float(len(self.funcs))
            self.consts[pdf.constants_idx + len(pdf.consts)] = pdf.consts[-1]

    def arrays(self):
        return (np.asarray(self.funcs, np.int32), np.asarray(self.params, np.float64),
                np.asarray(self.consts, np.float64), np.asarray(self.obs, np.float64),
                np.asarray(self.norms, np.float64))


def flatten(top: Pdf) -> Program:
    """setupObservables (top's list numbered 0..k-1) then setIndices."""
    for i, o in enumerate(top.observables):
        o.index = i
    prog = Program()
    prog.populate(top)
    return prog


class Gaussian(Pdf):
    func = FN["GAUSSIAN"]

    def __init__(self, name, x, mean, sigma):
        super().__init__(name, [x], [mean, sigma])

    def has_analytic_integral(self):
        return True

    def integrate(self, lo, hi):
        # GaussianPdf::integrate, src/PDFs/basic/GaussianPdf.cu:28-35 (over all R)
        root_pi = math.sqrt(math.atan2(0.0, -1.0))
        sigma = self.params[1].value
        sigma *= 1.4142135623730951 * root_pi
        return sigma


class Exp(Pdf):
    func = FN["EXP"]

    def __init__(self, name, x, alpha):
        super().__init__(name, [x], [alpha])

    def has_analytic_integral(self):
        return len(self.params) == 1

    def integrate(self, lo, hi):
        # ExpPdf::integrate, src/PDFs/basic/ExpPdf.cu:121-132
        alpha = self.params[0].value
        if alpha == 0:
            return hi - lo
        ret = math.exp(alpha * hi) - math.exp(alpha * lo)
        ret /= alpha
        return ret


class ExpGaus(Pdf):
    func = FN["EXPGAUS"]

    def __init__(self, name, x, mean, sigma, tau):
        super().__init__(name, [x], [mean, sigma, tau])


class MultiPolynomial(Pdf):
    """PolynomialPdf multivariate constructor, src/PDFs/basic/PolynomialPdf.cu:166-215."""
    func = FN["MULTI_POLYNOMIAL"]

    def __init__(self, name, obses, coeffs, offsets, max_degree):
        consts = [max_degree]
        nparams = 1
        for i in range(len(obses)):
            nparams *= (max_degree + 1 + i)
            consts.append(max_degree + 1 + i)
        for i in range(len(obses), 1, -1):
            nparams //= i
        coeffs = list(coeffs)
        while nparams > len(coeffs):
            coeffs.append(Var(f"{name}_extra_coeff_{len(coeffs)}", 0, fixed=True))
        offsets = list(offsets)
        while len(offsets) < len(obses):
            offsets.append(Var(f"{name}_extra_offset_{len(offsets)}", 0, fixed=True))
        super().__init__(name, list(obses), offsets + coeffs, consts)


class Add(Pdf):
    """AddPdf (src/PDFs/combine/AddPdf.cu:85-141)."""

    def __init__(self, name, weights, comps):
        super().__init__(name, [], list(weights), [], comps)
        self.extended = len(weights) == len(comps)
        self.func = FN["ADD_EXT"] if self.extended else FN["ADD"]
        self.observables = self.all_observables()


class Prod(Pdf):
    """ProdPdf (src/PDFs/combine/ProdPdf.cu:30-73); components on disjoint observables."""
    func = FN["PROD"]

    def __init__(self, name, comps):
        super().__init__(name, [], [], [len(comps)], comps)
        self.observables = self.all_observables()


class Resonance(Pdf):
    def __init__(self, name, func, ar, ai, params=(), consts=()):
        super().__init__(name, [], params, consts)
        self.func = func
        self.amp_real, self.amp_imag = ar, ai


def RBW(name, ar, ai, mass, width, spin, cyc, norm=True, sym=False):
    """Resonances::RBW (src/PDFs/physics/resonances/RBW.cu:93-117)."""
    return Resonance(name, FN["RBW_SYM"] if sym else FN["RBW"], ar, ai, [mass, width], [spin, cyc, 1 if norm else 0])


def GS(name, ar, ai, mass, width, spin, cyc, norm=True):
    """Resonances::GS (src/PDFs/physics/resonances/GS.cu:81-101)."""
    return Resonance(name, FN["GS"], ar, ai, [mass, width], [spin, cyc, 1 if norm else 0])


def NonRes(name, ar, ai):
    return Resonance(name, FN["NONRES"], ar, ai)


class Dalitz(Pdf):
    """Amp3Body (src/PDFs/physics/Amp3Body.cu:127-196)."""
    func = FN["DALITZ"]

    def __init__(self, name, m12, m13, evtnum, decay, resonances, eff, cache_to_use=0):
        params = []
        for r in resonances:
            params += [r.amp_real, r.amp_imag]
        super().__init__(name, [m12, m13, evtnum], params, [len(resonances), cache_to_use],
                         list(resonances) + [eff])
        self.decay = list(decay)  # motherMass, d1, d2, d3, meson_radius, mother_meson_radius
        self.m12, self.m13 = m12, m13
        self.resonances = list(resonances)
        self.cache_to_use = cache_to_use


class TruthResolution(Pdf):
    """TruthResolution (src/PDFs/physics/TruthResolution.cu:48-53): no parameters."""
    func = FN["TRUTH_RES"]

    def __init__(self, name="truth"):
        super().__init__(name)


class ThreeGaussResolution(Pdf):
    """ThreeGaussResolution (src/PDFs/physics/ThreeGaussResolution.cu:131-151): parameters
    core/tail fractions, core bias/scale, tail bias/scale, outlier bias/scale,
    selection bias, then the two common bias offsets and width scales."""
    func = FN["THREEGAUSS_RES"]

    def __init__(self, name, cf, tf, cb, cs, tb, ts, ob, os_, sb, cbo, cws, cbo2, cws2):
        super().__init__(name, [], [cf, tf, cb, cs, tb, ts, ob, os_, sb, cbo, cws, cbo2, cws2])
        self.selection_bias = sb


class Tddp(Pdf):
    """Amp3Body_TD (src/PDFs/physics/Amp3Body_TD.cu:262-359): parameters tau, x, y,
    dx, dy and the couplings; constants nres, cacheToUse, mistag flag, end of the
    efficiency subtree; observables dtime, sigma_t, m12, m13, event number
    [, mistag[, charm tag]].  mix = (tau, x, y, dx, dy) Vars."""
    func = FN["TDDP"]

    def __init__(self, name, dtime, sigmat, m12, m13, evtnum, decay, mix, resonances, resolution, eff,
                 mistag=None, charmtag=None, cache_to_use=0):
        params = list(mix)
        for r in resonances:
            params += [r.amp_real, r.amp_imag]
        obs = [dtime, sigmat, m12, m13, evtnum]
        if mistag is not None:
            obs.append(mistag)
        if charmtag is not None:
            obs.append(charmtag)
        super().__init__(name, obs, params, [len(resonances), cache_to_use, 0 if mistag is not None else 1, 0],
                         list(resonances) + [resolution, eff])
        self.decay = list(decay)
        self.m12, self.m13 = m12, m13
        self.resonances = list(resonances)
        self.resolution = resolution
        self.mistag = mistag
        self.cache_to_use = cache_to_use
        self.d0_fraction = 0.5


class OracleProg:
    """A gfo_prog (the oracle's flattened-tree evaluator) for one top-level PDF."""

    def __init__(self, top: Pdf, decay=None):
        self.top = top
        self.prog = flatten(top)
        f, p, c, o, n = self.prog.arrays()
        self.f, self.p, self.c, self.o, self.n = f, p, c, o, n
        d = np.asarray(decay if decay is not None else [0.0] * 6, np.float64)
        self.decay = d
        L = lib()
        self.h = L.gfo_prog_new(f.ctypes.data_as(_I), f.size, _dp(p), p.size, _dp(c), c.size, _dp(o), o.size,
                                _dp(n), n.size, _dp(d))
        self._caches = {}

    def __del__(self):
        try:
            lib().gfo_prog_free(self.h)
        except Exception:
            pass

    # -- parameter / norm sync --------------------------------------------
    def sync_params(self):
        vals = []
        for node in self.prog.nodes:
            vals.append(float(len(node.params)))
            vals += [q.value for q in node.params]
        self.p = np.asarray(vals, np.float64)
        lib().gfo_prog_set_params(self.h, _dp(self.p), self.p.size)

    def sync_norms(self):
        vals = []
        for node in self.prog.nodes:
            vals += [1.0, node.cached_norm]
        self.n = np.asarray(vals, np.float64)
        lib().gfo_prog_set_norms(self.h, _dp(self.n), self.n.size)

    def set_total_events(self, n):
        lib().gfo_prog_set_total_events(self.h, float(n))

    def set_cache(self, slot, waves):
        w = np.ascontiguousarray(waves.view(np.float64) if np.iscomplexobj(waves) else waves, dtype=np.float64)
        self._caches[slot] = w
        lib().gfo_prog_set_cache(self.h, slot, _dp(w), w.size // 2)

    def error(self):
        return lib().gfo_prog_error(self.h)

    # -- evaluation ---------------------------------------------------------
    def eval_event(self, evt):
        e = np.ascontiguousarray(evt, np.float64)
        return lib().gfo_eval_event(self.h, _dp(e))

    def eval_at(self, func_idx, evt):
        e = np.zeros(16)
        e[:len(evt)] = evt
        return lib().gfo_eval_event_at(self.h, func_idx, _dp(e))

    def eval_metric(self, aos, metric):
        aos = np.ascontiguousarray(aos, np.float64)
        n, k = aos.shape
        out = np.empty(n)
        lib().gfo_eval_metric(self.h, _dp(aos), n, k, METRIC[metric] if isinstance(metric, str) else metric, _dp(out))
        return out

    def reduce_metric(self, aos, metric="NLL", nchunks=None):
        aos = np.ascontiguousarray(aos, np.float64)
        n, k = aos.shape
        return lib().gfo_reduce_metric(self.h, _dp(aos), n, k, METRIC[metric] if isinstance(metric, str) else metric,
                                       nchunks or num_procs())

    def reduce_bins(self, func_idx, ranges, nchunks=None):
        r = np.ascontiguousarray(ranges, np.float64)
        return lib().gfo_reduce_bins(self.h, func_idx, _dp(r), r.size // 3, nchunks or num_procs())

    def eval_bins(self, func_idx, ranges):
        r = np.ascontiguousarray(ranges, np.float64)
        total = int(np.prod([int(math.floor(r[3 * i + 2] + 0.5)) for i in range(r.size // 3)]))
        out = np.empty(total)
        lib().gfo_eval_bins(self.h, func_idx, _dp(r), r.size // 3, _dp(out))
        return out

    # -- Dalitz ---------------------------------------------------------------
    def dalitz_layout(self, dalitz_func):
        res = np.zeros(64, np.int32)
        eff = np.zeros(1, np.int32)
        n = lib().gfo_dalitz_layout(self.h, dalitz_func, res.ctypes.data_as(_I), eff.ctypes.data_as(_I))
        return [int(x) for x in res[:n]], int(eff[0])

    def lineshape(self, res_func, m12, m13):
        out = np.zeros(2)
        lib().gfo_lineshape(self.h, res_func, m12, m13, _dp(out))
        return complex(out[0], out[1])

    def dalitz_wave(self, dalitz_func, res_func, aos, offset=0, n=None):
        aos = np.ascontiguousarray(aos, np.float64)
        rows, k = aos.shape
        n = rows - offset if n is None else n
        out = np.zeros(2 * n)
        lib().gfo_dalitz_wave(self.h, dalitz_func, res_func, _dp(aos), n, k, offset, _dp(out))
        return out.view(np.complex128)

    def set_cache_b(self, slot, waves):
        w = np.ascontiguousarray(waves.view(np.float64) if np.iscomplexobj(waves) else waves, dtype=np.float64)
        self._caches[(slot, "b")] = w
        lib().gfo_prog_set_cache_b(self.h, slot, _dp(w))

    def tddp_layout(self, tddp_func):
        res = np.zeros(64, np.int32)
        out = np.zeros(3, np.int32)
        n = lib().gfo_tddp_layout(self.h, tddp_func, res.ctypes.data_as(_I), out[0:].ctypes.data_as(_I),
                                  out[1:].ctypes.data_as(_I), out[2:].ctypes.data_as(_I))
        return [int(x) for x in res[:n]], int(out[0]), int(out[1]), int(out[2])

    def tddp_wave(self, tddp_func, res_func, aos, offset=0, n=None):
        aos = np.ascontiguousarray(aos, np.float64)
        rows, k = aos.shape
        n = rows - offset if n is None else n
        a = np.zeros(2 * n)
        b = np.zeros(2 * n)
        lib().gfo_tddp_wave(self.h, tddp_func, res_func, _dp(aos), n, k, offset, _dp(a), _dp(b))
        return a.view(np.complex128), b.view(np.complex128)

    def tddp_integrals(self, tddp_func, grid6, nchunks=None):
        res = self.tddp_layout(tddp_func)[0]
        g = np.asarray(grid6, np.float64)
        m = len(res)
        outs = [np.zeros(2 * m * m) for _ in range(3)]
        lib().gfo_tddp_integrals(self.h, tddp_func, _dp(g), nchunks or num_procs(), *(_dp(o) for o in outs))
        return tuple(o.view(np.complex128).reshape(m, m) for o in outs)

    def dalitz_integrals(self, dalitz_func, grid6, eff_func, nchunks=None):
        res, _ = self.dalitz_layout(dalitz_func)
        g = np.asarray(grid6, np.float64)
        out = np.zeros(2 * len(res) * len(res))
        lib().gfo_dalitz_integrals(self.h, dalitz_func, _dp(g), eff_func, nchunks or num_procs(), _dp(out))
        return out.view(np.complex128).reshape(len(res), len(res))


class RawProg(OracleProg):
    """A gfo_prog built directly from flattened arrays (e.g. those of the engine's API)."""

    def __init__(self, funcs, params, consts, obs, norms, decay=None):
        self.f = np.ascontiguousarray(funcs, np.int32)
        self.p, self.c, self.o, self.n = (np.ascontiguousarray(a, np.float64) for a in (params, consts, obs, norms))
        self.decay = np.asarray(decay if decay is not None else [0.0] * 6, np.float64)
        L = lib()
        self.h = L.gfo_prog_new(self.f.ctypes.data_as(_I), self.f.size, _dp(self.p), self.p.size, _dp(self.c),
                                self.c.size, _dp(self.o), self.o.size, _dp(self.n), self.n.size, _dp(self.decay))
        self._caches = {}

    def set_norms(self, norms):
        self.n = np.ascontiguousarray(norms, np.float64)
        lib().gfo_prog_set_norms(self.h, _dp(self.n), self.n.size)

    def set_params(self, params):
        self.p = np.ascontiguousarray(params, np.float64)
        lib().gfo_prog_set_params(self.h, _dp(self.p), self.p.size)


def resolution_norm(kind, di, tau, x, y, sel_bias=0.0):
    """MixingTimeResolution::normalization of kind FN["TRUTH_RES"] / FN["THREEGAUSS_RES"]."""
    return lib().gfo_resolution_norm(kind, *[float(v) for v in di], tau, x, y, sel_bias)


def resolution_eval(kind, params13, t1, t2, abr, abi, tau, dtime, x, y, sigma):
    """Resolution function value (device_truth_resolution / device_threegauss_resolution)."""
    p = np.ascontiguousarray(params13 if params13 is not None else np.zeros(13), np.float64)
    return lib().gfo_resolution_eval(kind, _dp(p), t1, t2, abr, abi, tau, dtime, x, y, sigma)


def tddp_norm_from_integrals(coups, Iaa, Iab, Ibb, mix5, res_kind, sel_bias, d0_fraction, mistag_on, mistag,
                             bs12, bs13):
    c = np.ascontiguousarray(np.asarray(coups, np.complex128).view(np.float64))
    arrs = [np.ascontiguousarray(np.asarray(I, np.complex128).view(np.float64)) for I in (Iaa, Iab, Ibb)]
    m = np.ascontiguousarray(mix5, np.float64)
    return lib().gfo_tddp_norm_from_integrals(len(coups), _dp(c), *(_dp(a) for a in arrs), _dp(m), res_kind,
                                              sel_bias, d0_fraction, int(bool(mistag_on)), mistag, bs12, bs13)


def dalitz_norm_from_integrals(coups, I, bs12, bs13):
    c = np.ascontiguousarray(np.asarray(coups, np.complex128).view(np.float64))
    Ia = np.ascontiguousarray(np.asarray(I, np.complex128).view(np.float64))
    return lib().gfo_dalitz_norm_from_integrals(len(coups), _dp(c), _dp(Ia), bs12, bs13)


# ----------------------------------------------------------------------------
# Host logic: normalize() and calculateNLL() restated over the Python model.
# ----------------------------------------------------------------------------

class Model:
    """Top-level PDF + data, mirroring GooPdf's calculateNLL / normalize."""

    def __init__(self, top: Pdf, aos=None, decay=None, nchunks=None):
        self.top = top
        self.op = OracleProg(top, decay)
        self.nchunks = nchunks or num_procs()
        self.aos = None
        self.integrals = {}
        self._reset_caches()
        if aos is not None:
            self.set_data(aos)

    def set_data(self, aos):
        self.aos = np.ascontiguousarray(aos, np.float64)
        self.op.set_total_events(self.aos.shape[0])
        self._reset_caches()

    def _reset_caches(self):
        for node in self.op.prog.nodes:
            if isinstance(node, (Dalitz, Tddp)):
                node.redo = [True] * len(node.resonances)
                node.cached_mw = [None] * len(node.resonances)
                node.waves = [None] * len(node.resonances)

    def normalize(self, pdf: Pdf = None):
        pdf = pdf or self.top
        if isinstance(pdf, Add):
            return self._normalize_add(pdf)
        if isinstance(pdf, Prod):
            # ProdPdf::normalize, ProdPdf.cu:75-99 (no overlapping observables)
            for c in pdf.components:
                self.normalize(c)
            pdf.cached_norm = 1.0
            return 1.0
        if isinstance(pdf, Dalitz):
            return self._normalize_dalitz(pdf)
        if isinstance(pdf, Tddp):
            return self._normalize_tddp(pdf)
        return self._normalize_goopdf(pdf)

    def _normalize_goopdf(self, pdf):
        # GooPdf::normalize, src/PDFs/GooPdf.cu:340-396
        ret = 1.0
        if pdf.has_analytic_integral():
            for o in pdf.observables:
                ret *= pdf.integrate(o.lo, o.hi)
            pdf.cached_norm = 1.0 / ret
            return ret
        ranges, total, vol = [], 1, 1.0
        for o in pdf.observables:
            nb = pdf.integration_bins if pdf.integration_bins > 0 else o.nbins
            ranges += [o.lo, o.hi, nb]
            total *= nb
            vol *= o.hi - o.lo
        ret *= vol
        ret /= total
        self.op.sync_params()
        s = self.op.reduce_bins(pdf.function_idx, ranges, self.nchunks)
        if math.isnan(s) or s <= 0:
            raise RuntimeError(f"{pdf.name}: bad normalisation sum {s}")
        ret *= s
        pdf.cached_norm = 1.0 / ret
        return ret

    def _normalize_add(self, pdf):
        # AddPdf::normalize, src/PDFs/combine/AddPdf.cu:143-185
        ret, tw = 0.0, 0.0
        for i in range(len(pdf.components) - 1):
            w = pdf.params[i].value
            tw += w
            curr = self.normalize(pdf.components[i])
            ret += curr * w
        last = self.normalize(pdf.components[-1])
        if pdf.extended:
            lw = pdf.params[len(pdf.components) - 1].value
            tw += lw
            ret += last * lw
            ret /= tw
        else:
            ret += (1 - tw) * last
        pdf.cached_norm = 1.0
        return ret

    def _normalize_dalitz(self, pdf: Dalitz):
        # Amp3Body::normalize, src/PDFs/physics/Amp3Body.cu:238-359
        def set1(node):
            node.cached_norm = 1.0
            for c in node.components:
                set1(c)
        set1(pdf)
        self.op.sync_params()
        self.op.sync_norms()
        nres = len(pdf.resonances)
        res_funcs, eff_func = self.op.dalitz_layout(pdf.function_idx)
        for i, r in enumerate(pdf.resonances):
            mw = tuple(q.value for q in r.params)
            if pdf.cached_mw[i] != mw:
                pdf.redo[i] = True
                pdf.cached_mw[i] = mw
        if any(pdf.redo):
            if self.aos is not None:
                for i in range(nres):
                    if pdf.redo[i]:
                        pdf.waves[i] = self.op.dalitz_wave(pdf.function_idx, res_funcs[i], self.aos)
                        self.op.set_cache(16 * pdf.cache_to_use + i, pdf.waves[i])
            grid6 = [pdf.m12.lo, pdf.m12.hi, pdf.m12.nbins, pdf.m13.lo, pdf.m13.hi, pdf.m13.nbins]
            self.integrals[pdf.function_idx] = self.op.dalitz_integrals(pdf.function_idx, grid6, eff_func,
                                                                        self.nchunks)
            pdf.redo = [False] * nres
        I = self.integrals[pdf.function_idx]
        coups = [complex(r.amp_real.value, r.amp_imag.value) for r in pdf.resonances]
        ret = dalitz_norm_from_integrals(coups, I, pdf.m12.bin_size(), pdf.m13.bin_size())
        pdf.cached_norm = 1.0 / ret
        return ret

    def _normalize_tddp(self, pdf: Tddp):
        # Amp3Body_TD::normalize, src/PDFs/physics/Amp3Body_TD.cu:590-748
        def set1(node):
            node.cached_norm = 1.0
            for c in node.components:
                set1(c)
        set1(pdf)
        self.op.sync_params()
        self.op.sync_norms()
        nres = len(pdf.resonances)
        res_funcs, _, _, _ = self.op.tddp_layout(pdf.function_idx)
        for i, r in enumerate(pdf.resonances):
            mw = tuple(q.value for q in r.params)
            if pdf.cached_mw[i] != mw:
                pdf.redo[i] = True
                pdf.cached_mw[i] = mw
        if any(pdf.redo):
            if self.aos is not None:
                for i in range(nres):
                    if pdf.redo[i]:
                        a, b = self.op.tddp_wave(pdf.function_idx, res_funcs[i], self.aos)
                        pdf.waves[i] = (a, b)
                        self.op.set_cache(16 * pdf.cache_to_use + i, a)
                        self.op.set_cache_b(16 * pdf.cache_to_use + i, b)
            grid6 = [pdf.m12.lo, pdf.m12.hi, pdf.m12.nbins, pdf.m13.lo, pdf.m13.hi, pdf.m13.nbins]
            self.integrals[pdf.function_idx] = self.op.tddp_integrals(pdf.function_idx, grid6, self.nchunks)
            pdf.redo = [False] * nres
        Iaa, Iab, Ibb = self.integrals[pdf.function_idx]
        coups = [complex(r.amp_real.value, r.amp_imag.value) for r in pdf.resonances]
        mix5 = [q.value for q in pdf.params[:5]]
        res = pdf.resolution
        sel = res.selection_bias.value if isinstance(res, ThreeGaussResolution) else 0.0
        mistag_on = False
        if pdf.mistag is not None:
            # Observable::operator bool (include/goofit/Variable.h:103-105)
            v, lo, hi = pdf.mistag.value, pdf.mistag.lo, pdf.mistag.hi
            mistag_on = hi == lo or (lo <= v <= hi)
        ret = tddp_norm_from_integrals(coups, Iaa, Iab, Ibb, mix5, res.func, sel, pdf.d0_fraction, mistag_on,
                                       pdf.mistag.value if pdf.mistag is not None else 0.0, pdf.m12.bin_size(),
                                       pdf.m13.bin_size())
        pdf.cached_norm = 1.0 / ret
        return ret

    def calculate_nll(self):
        # GooPdf::calculateNLL, src/PDFs/GooPdf.cu:254-273 (+ AddPdf extended term, AddPdf.cu:187-204)
        self.normalize()
        self.op.sync_params()
        self.op.sync_norms()
        if self.top.cached_norm <= 0:
            raise RuntimeError("non-positive normalization")
        ret = self.op.reduce_metric(self.aos, "NLL", self.nchunks)
        if ret == 0.0:
            raise RuntimeError("zero NLL")
        if isinstance(self.top, Add) and self.top.extended:
            exp_events = sum(p.value for p in self.top.params)
            ret += exp_events - self.aos.shape[0] * math.log(exp_events)
        return ret

    def nll_terms(self):
        self.normalize()
        self.op.sync_params()
        self.op.sync_norms()
        return self.op.eval_metric(self.aos, "NLL")
