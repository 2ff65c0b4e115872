"""ctypes binding of the C ABI in include/goofit_hip.h (libgoofit_hip.so).

This is the thinnest possible host path onto the gfx950 engine: every method
is one C call.  Loading fails loudly when the library has not been built; there
is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libgoofit_hip.so")

# include/goofit_hip.h
GF_OK, GF_ERR_INVALID, GF_ERR_HIP, GF_ERR_NOMEM, GF_ERR_STATE, GF_ERR_NODEVICE = range(6)
(TUNE_INLINE_ARGS, TUNE_FAST_DALITZ, TUNE_BLOCKS_PER_CU, TUNE_FAST_BLOCKS_PER_CU, TUNE_SPIN_US,
 TUNE_ZERO_OUTSIDE, TUNE_FLAT, TUNE_FUSED_FINISH, TUNE_COH_MFMA, TUNE_BATCH_MX_PTS) = range(10)
FN = dict(NONE=0, GAUSSIAN=1, EXP=2, EXP_OFFSET=3, EXP_POLY=4, EXP_POLY_OFFSET=5, EXPGAUS=6,
          POLYNOMIAL=7, OFFSET_POLYNOMIAL=8, MULTI_POLYNOMIAL=9, ADD=10, ADD_EXT=11, PROD=12,
          DALITZ=13, RBW=14, RBW_SYM=15, GS=16, NONRES=17)
METRIC = dict(EVAL=0, NLL=1, PROB=2, BINAVG=3, BINWITHERROR=4, CHISQ=5)

# Every symbol the header declares (checked by the tests against the built library).
EXPORTS = [
    "gf_device_count", "gf_ctx_create", "gf_ctx_destroy", "gf_last_error", "gf_ctx_set_tuning",
    "gf_ctx_get_tuning",
    "gf_events_upload", "gf_events_upload_aos", "gf_events_alloc", "gf_events_column",
    "gf_events_set_identity_column", "gf_program_set", "gf_params_update", "gf_norms_update",
    "gf_decay3_set", "gf_reduce_metric", "gf_reduce_metric_noalloc", "gf_reduce_metric_multi", "gf_eval_metric", "gf_eval_metric_device", "gf_reduce_bins",
    "gf_eval_bins", "gf_grid_cells", "gf_dalitz_grid_cells",
    "gf_dalitz_set_data_size", "gf_dalitz_cache", "gf_dalitz_gram", "gf_dalitz_gram_points",
    "gf_dalitz_gram_points_device", "gf_dalitz_get_cache", "gf_dalitz_cache_gram", "gf_tddp_gram",
    "gf_time_reduce_metric", "gf_time_reduce_sweep", "gf_time_sweep_back_to_back", "gf_time_dalitz_cache",
    "gf_time_dalitz_gram",
    "gf_sample_uniform", "gf_mlp_param_count", "gf_mlp_create", "gf_mlp_destroy", "gf_mlp_get_params",
    "gf_mlp_set_params", "gf_mlp_forward", "gf_mlp_gradient", "gf_mlp_train", "gf_mlp_build_proposal",
    "gf_mlp_sample", "gf_mlp_train_host", "gf_mlp_use_graph", "gf_dalitz_set_norm_points", "gf_dalitz_gram_norm_points",
    "gf_dalitz_generate", "gf_batch_create", "gf_batch_destroy", "gf_batch_set_toy", "gf_batch_reduce",
    "gf_batch_set_chunk", "gf_batch_get_chunk", "gf_batch_use_graph", "gf_batch_autotune",
    "gf_batch_add_point", "gf_batch_reduce_points", "gf_batch_clear_points", "gf_batch_submit_points",
    "gf_batch_wait_points", "gf_batch_set_norm_ops", "gf_dalitz_set_norm_rule", "gf_mlp_train_generated",
]

_D = ctypes.POINTER(ctypes.c_double)
# gf_norm_op of include/goofit_hip.h
NORM_OP_DTYPE = np.dtype([("kind", "<i4"), ("out", "<i4"), ("p", "<i4"), ("n", "<i4"), ("g", "<i4"), ("pad", "<i4"),
                          ("c0", "<f8"), ("c1", "<f8")])
_I = ctypes.POINTER(ctypes.c_int32)
_lib = None


class GooFitError(RuntimeError):
    """A failing C-ABI call (status code + gf_last_error text)."""

    def __init__(self, status, msg):
        super().__init__(f"[{status}] {msg}")
        self.status = status


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run `python -m goofit_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        c_p = ctypes.c_void_p
        L.gf_device_count.restype = ctypes.c_int
        L.gf_last_error.restype = ctypes.c_char_p
        L.gf_ctx_create.argtypes = [ctypes.POINTER(c_p), ctypes.c_int]
        L.gf_ctx_destroy.argtypes = [c_p]
        L.gf_ctx_destroy.restype = None
        L.gf_ctx_set_tuning.argtypes = [c_p, ctypes.c_int, ctypes.c_int]
        L.gf_ctx_get_tuning.argtypes = [c_p, ctypes.c_int, _I]
        L.gf_events_upload.argtypes = [c_p, ctypes.c_int, ctypes.c_int64, ctypes.POINTER(_D)]
        L.gf_events_upload_aos.argtypes = [c_p, ctypes.c_int, ctypes.c_int64, _D]
        L.gf_events_alloc.argtypes = [c_p, ctypes.c_int, ctypes.c_int64]
        L.gf_events_column.argtypes = [c_p, ctypes.c_int, ctypes.POINTER(c_p)]
        L.gf_events_set_identity_column.argtypes = [c_p, ctypes.c_int]
        L.gf_program_set.argtypes = [c_p, _I, ctypes.c_int, _D, ctypes.c_int, _D, ctypes.c_int, _D, ctypes.c_int,
                                     _D, ctypes.c_int]
        L.gf_params_update.argtypes = [c_p, _D, ctypes.c_int]
        L.gf_norms_update.argtypes = [c_p, _D, ctypes.c_int]
        L.gf_decay3_set.argtypes = [c_p, _D]
        L.gf_reduce_metric.argtypes = [c_p, ctypes.c_int, _D, _D]
        L.gf_eval_metric.argtypes = [c_p, ctypes.c_int, _D, ctypes.c_int64]
        L.gf_eval_metric_device.argtypes = [c_p, ctypes.c_int, c_p, ctypes.c_int64]
        L.gf_reduce_bins.argtypes = [c_p, ctypes.c_int, _D, ctypes.c_int, _D]
        for name, args in (("gf_eval_bins", [c_p, ctypes.c_int, _D, ctypes.c_int, _D, ctypes.c_int64]),
                           ("gf_grid_cells", [c_p, _D, ctypes.c_int, _D, ctypes.c_int64]),
                           ("gf_dalitz_grid_cells", [c_p, _D, ctypes.c_int, _D, _D, ctypes.c_int64])):
            if hasattr(L, name):  # a library built before these entry points fails only their callers
                getattr(L, name).argtypes = args
        L.gf_dalitz_set_data_size.argtypes = [c_p, ctypes.c_int, ctypes.c_int64, ctypes.c_int64]
        L.gf_dalitz_cache.argtypes = [c_p, ctypes.c_int, ctypes.c_uint64]
        L.gf_dalitz_gram.argtypes = [c_p, ctypes.c_int, _D, ctypes.c_int, _D]
        L.gf_dalitz_gram_points.argtypes = [c_p, ctypes.c_int, _D, _D, _D, ctypes.c_int64, ctypes.c_int, _D]
        L.gf_dalitz_gram_points_device.argtypes = [c_p, ctypes.c_int, c_p, c_p, c_p, ctypes.c_int64, ctypes.c_int,
                                                   _D]
        L.gf_dalitz_get_cache.argtypes = [c_p, ctypes.c_int, ctypes.c_int, _D, ctypes.c_int64]
        L.gf_dalitz_cache_gram.argtypes = [c_p, ctypes.c_int, _D]
        L.gf_tddp_gram.argtypes = [c_p, ctypes.c_int, _D, _D, _D, _D]
        L.gf_reduce_metric_multi.argtypes = [c_p, ctypes.c_int, ctypes.c_int, _D, ctypes.c_int, _D, ctypes.c_int, _D]
        if hasattr(L, "gf_reduce_metric_noalloc"):
            L.gf_reduce_metric_noalloc.argtypes = [c_p, ctypes.c_int, c_p, c_p]
        L.gf_time_reduce_metric.argtypes = [c_p, ctypes.c_int, ctypes.c_int, _D]
        L.gf_time_reduce_sweep.argtypes = [c_p, ctypes.c_int, ctypes.c_int, _D, _D]
        L.gf_time_sweep_back_to_back.argtypes = [c_p, ctypes.c_int, ctypes.c_int, _D]
        _F = ctypes.POINTER(ctypes.c_float)
        L.gf_sample_uniform.argtypes = [c_p, _D, ctypes.c_int64, ctypes.c_uint64, c_p, c_p, c_p]
        L.gf_mlp_param_count.restype = ctypes.c_int
        L.gf_mlp_create.argtypes = [c_p, ctypes.c_uint64, ctypes.c_int, ctypes.POINTER(c_p)]
        L.gf_mlp_destroy.argtypes = [c_p]
        L.gf_mlp_destroy.restype = None
        L.gf_mlp_get_params.argtypes = [c_p, _F, ctypes.c_int]
        L.gf_mlp_set_params.argtypes = [c_p, _F, ctypes.c_int]
        L.gf_mlp_forward.argtypes = [c_p, _F, ctypes.c_int64, _F]
        L.gf_mlp_gradient.argtypes = [c_p, _F, _F, ctypes.c_int64, _F]
        L.gf_mlp_train.argtypes = [c_p, c_p, c_p, ctypes.c_int64, _D, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                   ctypes.c_uint64, _F, _F]
        L.gf_mlp_build_proposal.argtypes = [c_p, _D, ctypes.c_int, ctypes.c_double]
        L.gf_mlp_sample.argtypes = [c_p, ctypes.c_int64, ctypes.c_uint64, c_p, c_p, c_p]
        L.gf_mlp_use_graph.argtypes = [c_p, ctypes.c_int]
        L.gf_mlp_train_host.argtypes = [c_p, _D, _D, ctypes.c_int64, _D, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                        ctypes.c_uint64, _F, _F]
        L.gf_dalitz_set_norm_points.argtypes = [c_p, ctypes.c_int, c_p, ctypes.c_int64, ctypes.c_uint64, _D]
        L.gf_dalitz_gram_norm_points.argtypes = [c_p, ctypes.c_int, ctypes.c_int, ctypes.c_int64, ctypes.c_int64, _D]
        _L = ctypes.POINTER(ctypes.c_int64)
        L.gf_batch_create.argtypes = [c_p, ctypes.c_int, _L, _L, ctypes.POINTER(c_p)]
        L.gf_batch_destroy.argtypes = [c_p]
        L.gf_batch_destroy.restype = None
        L.gf_batch_set_toy.argtypes = [c_p, ctypes.c_int, _D, _D]
        L.gf_batch_reduce.argtypes = [c_p, ctypes.POINTER(ctypes.c_uint8), ctypes.c_int, _D]
        L.gf_batch_set_chunk.argtypes = [c_p, ctypes.c_int]
        L.gf_batch_get_chunk.argtypes = [c_p, _I]
        L.gf_batch_use_graph.argtypes = [c_p, ctypes.c_int]
        L.gf_batch_autotune.argtypes = [c_p, ctypes.c_int, _I, _D]
        L.gf_batch_add_point.argtypes = [c_p, ctypes.c_int, _D, _D]
        L.gf_batch_reduce_points.argtypes = [c_p, ctypes.c_int, _D]
        L.gf_batch_clear_points.argtypes = [c_p]
        L.gf_batch_submit_points.argtypes = [c_p, ctypes.c_int, ctypes.POINTER(ctypes.c_uint32)]
        L.gf_batch_wait_points.argtypes = [c_p, ctypes.c_uint32, _D]
        if hasattr(L, "gf_batch_set_norm_ops"):
            L.gf_batch_set_norm_ops.argtypes = [c_p, c_p, ctypes.c_int, _D, _D, ctypes.c_int]
        L.gf_dalitz_generate.argtypes = [c_p, ctypes.c_int, ctypes.c_int64, ctypes.c_uint64, _D, ctypes.c_double,
                                         ctypes.c_int, c_p, c_p, ctypes.POINTER(ctypes.c_int64), _D]
        L.gf_time_dalitz_cache.argtypes = [c_p, ctypes.c_int, ctypes.c_int, _D]
        L.gf_time_dalitz_gram.argtypes = [c_p, ctypes.c_int, _D, ctypes.c_int, ctypes.c_int, _D]
        _lib = L
    return _lib


def _check(status):
    if status != GF_OK:
        raise GooFitError(status, lib().gf_last_error().decode())


def _dp(a):
    return a.ctypes.data_as(_D)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def device_count() -> int:
    return lib().gf_device_count()


class Context:
    """One gf_ctx: a HIP stream and the device state of one flattened PDF tree."""

    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        _check(lib().gf_ctx_create(ctypes.byref(h), device))
        self.h = h
        self.nevents = 0
        self._keep = []

    def set_tuning(self, key: int, value: int):
        _check(lib().gf_ctx_set_tuning(self.h, key, value))

    def get_tuning(self, key: int) -> int:
        v = ctypes.c_int32()
        _check(lib().gf_ctx_get_tuning(self.h, key, ctypes.byref(v)))
        return v.value

    def close(self):
        if self.h:
            lib().gf_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- events -----------------------------------------------------------
    def upload_columns(self, cols):
        cols = [_f64(c) for c in cols]
        n = cols[0].size if cols else 0
        arr = (_D * len(cols))(*[_dp(c) for c in cols])
        _check(lib().gf_events_upload(self.h, len(cols), n, arr))
        self.nevents = n

    def upload_aos(self, aos):
        aos = _f64(aos)
        _check(lib().gf_events_upload_aos(self.h, aos.shape[1], aos.shape[0], _dp(aos)))
        self.nevents = aos.shape[0]

    def alloc_events(self, ncols, n):
        _check(lib().gf_events_alloc(self.h, ncols, n))
        self.nevents = n

    def column_ptr(self, k) -> int:
        p = ctypes.c_void_p()
        _check(lib().gf_events_column(self.h, k, ctypes.byref(p)))
        return p.value

    def set_identity_column(self, k):
        _check(lib().gf_events_set_identity_column(self.h, k))

    # -- program ----------------------------------------------------------
    def set_program(self, funcs, params, consts, obs, norms):
        f = np.ascontiguousarray(funcs, dtype=np.int32)
        p, c, o, n = _f64(params), _f64(consts), _f64(obs), _f64(norms)
        _check(lib().gf_program_set(self.h, f.ctypes.data_as(_I), f.size, _dp(p), p.size, _dp(c), c.size, _dp(o),
                                    o.size, _dp(n), n.size))

    def update_params(self, params):
        p = _f64(params)
        _check(lib().gf_params_update(self.h, _dp(p), p.size))

    def update_norms(self, norms):
        n = _f64(norms)
        _check(lib().gf_norms_update(self.h, _dp(n), n.size))

    def set_decay(self, decay6):
        d = _f64(decay6)
        _check(lib().gf_decay3_set(self.h, _dp(d)))

    # -- Monte-Carlo normalisation points ---------------------------------------
    def dalitz_set_norm_points(self, dalitz_func, n, seed, box=None, proposal=None):
        b = np.ascontiguousarray(box if box is not None else [0, 0, 0, 0], np.float64)
        _check(lib().gf_dalitz_set_norm_points(self.h, dalitz_func, proposal.h if proposal is not None else None,
                                               n, seed, _dp(b)))

    def dalitz_gram_norm_points(self, dalitz_func, eff_func, offset, count, nres):
        out = np.zeros(2 * nres * nres)
        _check(lib().gf_dalitz_gram_norm_points(self.h, dalitz_func, eff_func, offset, count, _dp(out)))
        return out.view(np.complex128).reshape(nres, nres)

    def dalitz_generate(self, dalitz_func, n, seed, box, fmax=0.0):
        """n events (host arrays m12, m13), the number of candidates tried and the envelope."""
        b = np.ascontiguousarray(box, np.float64)
        a12, a13 = np.empty(n), np.empty(n)
        tried, fm = ctypes.c_int64(), ctypes.c_double()
        _check(lib().gf_dalitz_generate(self.h, dalitz_func, n, seed, _dp(b), fmax, 1, a12.ctypes.data,
                                        a13.ctypes.data, ctypes.byref(tried), ctypes.byref(fm)))
        return a12, a13, tried.value, fm.value

    # -- reductions -------------------------------------------------------
    def reduce_metric(self, metric="NLL"):
        s, c = ctypes.c_double(), ctypes.c_double()
        m = METRIC[metric] if isinstance(metric, str) else metric
        _check(lib().gf_reduce_metric(self.h, m, ctypes.byref(s), ctypes.byref(c)))
        return s.value, c.value

    def reduce_metric_noalloc(self, dev_out: int, metric="NLL", stream: int = 0):
        """Enqueue the reduction without allocating: the 32-byte result block
        (sum, residual, error word, sequence number) goes to device address dev_out,
        on `stream` (a hipStream_t handle; 0: the context's stream).  Returns at once."""
        m = METRIC[metric] if isinstance(metric, str) else metric
        _check(lib().gf_reduce_metric_noalloc(self.h, m, ctypes.c_void_p(dev_out), ctypes.c_void_p(stream or None)))

    def reduce_metric_multi(self, metric, params_list, norms_list):
        """Metric sums for several (params, norms) sets over the same events."""
        P = np.ascontiguousarray(np.asarray(params_list, np.float64))
        N = np.ascontiguousarray(np.asarray(norms_list, np.float64))
        k = P.shape[0]
        out = np.zeros(k)
        m = METRIC[metric] if isinstance(metric, str) else metric
        _check(lib().gf_reduce_metric_multi(self.h, m, k, _dp(P), P.shape[1], _dp(N), N.shape[1], _dp(out)))
        return out

    def eval_metric(self, metric="EVAL"):
        out = np.empty(self.nevents)
        m = METRIC[metric] if isinstance(metric, str) else metric
        _check(lib().gf_eval_metric(self.h, m, _dp(out), out.size))
        return out

    def eval_metric_device(self, metric, dev_ptr, n):
        m = METRIC[metric] if isinstance(metric, str) else metric
        _check(lib().gf_eval_metric_device(self.h, m, ctypes.c_void_p(dev_ptr), n))

    def reduce_bins(self, func_idx, ranges):
        r = _f64(ranges)
        s = ctypes.c_double()
        _check(lib().gf_reduce_bins(self.h, func_idx, _dp(r), r.size // 3, ctypes.byref(s)))
        return s.value

    @staticmethod
    def _grid_bins(r):
        return int(np.prod([int(math.floor(r[3 * i + 2] + 0.5)) for i in range(r.size // 3)]))

    def eval_bins(self, func_idx, ranges):
        """Per-bin raw values of function func_idx on the grid (gf_eval_bins)."""
        r = _f64(ranges)
        out = np.empty(self._grid_bins(r))
        _check(lib().gf_eval_bins(self.h, func_idx, _dp(r), r.size // 3, _dp(out), out.size))
        return out

    def grid_cells(self, ranges):
        """Cell centres of the grid, shape (k, nbins) (gf_grid_cells)."""
        r = _f64(ranges)
        k, n = r.size // 3, self._grid_bins(r)
        out = np.empty(k * n)
        _check(lib().gf_grid_cells(self.h, _dp(r), k, _dp(out), n))
        return out.reshape(k, n)

    def dalitz_grid_cells(self, grid6, time_dependent=False):
        """Centres (m12, m13) of the integration cells of a Dalitz grid (gf_dalitz_grid_cells)."""
        g = _f64(grid6)
        n = int(math.floor(g[2] + 0.5)) * int(math.floor(g[5] + 0.5))
        m12, m13 = np.empty(n), np.empty(n)
        _check(lib().gf_dalitz_grid_cells(self.h, _dp(g), 1 if time_dependent else 0, _dp(m12), _dp(m13), n))
        return m12, m13

    # -- Dalitz -----------------------------------------------------------
    def dalitz_set_data_size(self, dalitz_func, n, offset=0):
        _check(lib().gf_dalitz_set_data_size(self.h, dalitz_func, n, offset))

    def dalitz_cache(self, dalitz_func, mask=(1 << 64) - 1):
        _check(lib().gf_dalitz_cache(self.h, dalitz_func, mask))

    def dalitz_gram(self, dalitz_func, grid6, eff_func, nres):
        g = _f64(grid6)
        out = np.zeros(2 * nres * nres)
        _check(lib().gf_dalitz_gram(self.h, dalitz_func, _dp(g), eff_func, _dp(out)))
        return out.view(np.complex128).reshape(nres, nres)

    def dalitz_gram_points(self, dalitz_func, m12, m13, w, eff_func, nres):
        a, b = _f64(m12), _f64(m13)
        wk = _f64(w) if w is not None else None
        out = np.zeros(2 * nres * nres)
        _check(lib().gf_dalitz_gram_points(self.h, dalitz_func, _dp(a), _dp(b), _dp(wk) if wk is not None else None,
                                           a.size, eff_func, _dp(out)))
        return out.view(np.complex128).reshape(nres, nres)

    def dalitz_gram_points_device(self, dalitz_func, m12_ptr, m13_ptr, w_ptr, npts, eff_func, nres):
        out = np.zeros(2 * nres * nres)
        _check(lib().gf_dalitz_gram_points_device(self.h, dalitz_func, ctypes.c_void_p(m12_ptr),
                                                  ctypes.c_void_p(m13_ptr), ctypes.c_void_p(w_ptr or 0), npts,
                                                  eff_func, _dp(out)))
        return out.view(np.complex128).reshape(nres, nres)

    def dalitz_get_cache(self, dalitz_func, res, n):
        out = np.zeros(2 * n)
        _check(lib().gf_dalitz_get_cache(self.h, dalitz_func, res, _dp(out), n))
        return out.view(np.complex128)

    def tddp_gram(self, tddp_func, grid6, nres):
        """ThreeComplex integrals (sum eff a_i a_j*, a_i b_j*, b_i b_j*) of a time-dependent node."""
        g = np.ascontiguousarray(grid6, np.float64)
        outs = [np.zeros(2 * nres * nres) for _ in range(3)]
        _check(lib().gf_tddp_gram(self.h, tddp_func, _dp(g), *(_dp(o) for o in outs)))
        return tuple(o.view(np.complex128).reshape(nres, nres) for o in outs)

    def dalitz_cache_gram(self, dalitz_func, nres):
        out = np.zeros(2 * nres * nres)
        _check(lib().gf_dalitz_cache_gram(self.h, dalitz_func, _dp(out)))
        return out.view(np.complex128).reshape(nres, nres)

    # -- timing -----------------------------------------------------------
    def time_reduce_metric(self, metric, reps):
        ms = ctypes.c_double()
        m = METRIC[metric] if isinstance(metric, str) else metric
        _check(lib().gf_time_reduce_metric(self.h, m, reps, ctypes.byref(ms)))
        return ms.value

    def time_reduce_sweep(self, metric, reps):
        """(summed device ms of the event sweeps alone, whole span ms)."""
        sw, tot = ctypes.c_double(), ctypes.c_double()
        m = METRIC[metric] if isinstance(metric, str) else metric
        _check(lib().gf_time_reduce_sweep(self.h, m, reps, ctypes.byref(sw), ctypes.byref(tot)))
        return sw.value, tot.value

    def time_sweep_back_to_back(self, metric, reps):
        """Device ms per event sweep, reps sweeps launched back to back between two events."""
        ms = ctypes.c_double()
        m = METRIC[metric] if isinstance(metric, str) else metric
        _check(lib().gf_time_sweep_back_to_back(self.h, m, reps, ctypes.byref(ms)))
        return ms.value

    def time_dalitz_cache(self, dalitz_func, reps):
        ms = ctypes.c_double()
        _check(lib().gf_time_dalitz_cache(self.h, dalitz_func, reps, ctypes.byref(ms)))
        return ms.value

    def time_dalitz_gram(self, dalitz_func, grid6, eff_func, reps):
        ms = ctypes.c_double()
        g = _f64(grid6)
        _check(lib().gf_time_dalitz_gram(self.h, dalitz_func, _dp(g), eff_func, reps, ctypes.byref(ms)))
        return ms.value


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class Proposal:
    """A gf_mlp: the learned importance-sampling proposal of a Context."""

    def __init__(self, ctx: Context, seed: int = 1, batch_cap: int = 65536):
        h = ctypes.c_void_p()
        _check(lib().gf_mlp_create(ctx.h, seed, batch_cap, ctypes.byref(h)))
        self.h, self.ctx = h, ctx
        self.nparam = lib().gf_mlp_param_count()

    def close(self):
        if self.h:
            lib().gf_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self):
        p = np.empty(self.nparam, np.float32)
        _check(lib().gf_mlp_get_params(self.h, _fp(p), p.size))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        _check(lib().gf_mlp_set_params(self.h, _fp(p), p.size))

    def forward(self, xy):
        xy = np.ascontiguousarray(xy, np.float32).reshape(-1, 2)
        out = np.empty(xy.shape[0], np.float32)
        _check(lib().gf_mlp_forward(self.h, _fp(xy), xy.shape[0], _fp(out)))
        return out

    def gradient(self, xy, coef):
        xy = np.ascontiguousarray(xy, np.float32).reshape(-1, 2)
        coef = np.ascontiguousarray(coef, np.float32)
        g = np.empty(self.nparam, np.float32)
        _check(lib().gf_mlp_gradient(self.h, _fp(xy), _fp(coef), xy.shape[0], _fp(g)))
        return g

    def train(self, dev_m12: int, dev_m13: int, n: int, box, steps, batch, lr, seed):
        b = np.ascontiguousarray(box, np.float64)
        l0, l1 = ctypes.c_float(), ctypes.c_float()
        _check(lib().gf_mlp_train(self.h, dev_m12, dev_m13, n, _dp(b), steps, batch, lr, seed, ctypes.byref(l0),
                                  ctypes.byref(l1)))
        return l0.value, l1.value

    def use_graph(self, on):
        _check(lib().gf_mlp_use_graph(self.h, 1 if on else 0))

    def train_host(self, m12, m13, box, steps, batch, lr, seed):
        a, b = _f64(m12), _f64(m13)
        bx = np.ascontiguousarray(box, np.float64)
        l0, l1 = ctypes.c_float(), ctypes.c_float()
        _check(lib().gf_mlp_train_host(self.h, _dp(a), _dp(b), a.size, _dp(bx), steps, batch, lr, seed,
                                       ctypes.byref(l0), ctypes.byref(l1)))
        return l0.value, l1.value

    def build(self, box, grid=512, alpha=0.1):
        b = np.ascontiguousarray(box, np.float64)
        _check(lib().gf_mlp_build_proposal(self.h, _dp(b), grid, alpha))

    def sample(self, n, seed, dev_m12: int, dev_m13: int, dev_w: int):
        _check(lib().gf_mlp_sample(self.h, n, seed, dev_m12, dev_m13, dev_w))


def sample_uniform(ctx: Context, box, n, seed, dev_m12: int, dev_m13: int, dev_w: int):
    b = np.ascontiguousarray(box, np.float64)
    _check(lib().gf_sample_uniform(ctx.h, _dp(b), n, seed, dev_m12, dev_m13, dev_w))


BATCH_MAX_POINTS = 64  # GF_BATCH_MAX_POINTS


class Batch:
    """A gf_batch: toys over row ranges of a Context's events, reduced in one launch."""

    def __init__(self, ctx: Context, first, count):
        f = np.ascontiguousarray(first, np.int64)
        c = np.ascontiguousarray(count, np.int64)
        h = ctypes.c_void_p()
        L64 = ctypes.POINTER(ctypes.c_int64)
        _check(lib().gf_batch_create(ctx.h, f.size, f.ctypes.data_as(L64), c.ctypes.data_as(L64), ctypes.byref(h)))
        self.h, self.n = h, f.size

    def close(self):
        if self.h:
            lib().gf_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_toy(self, t, params=None, norms=None):
        p = _f64(params) if params is not None else None
        n = _f64(norms) if norms is not None else None
        _check(lib().gf_batch_set_toy(self.h, t, _dp(p) if p is not None else None, _dp(n) if n is not None else None))

    def reduce(self, active=None, metric="NLL"):
        a = np.ones(self.n, np.uint8) if active is None else np.ascontiguousarray(active, np.uint8)
        out = np.full(self.n, np.nan)
        m = METRIC[metric] if isinstance(metric, str) else metric
        _check(lib().gf_batch_reduce(self.h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), m, _dp(out)))
        return out

    def add_point(self, t, params, norms=None):
        """norms None: computed in the launch by the installed normalisation ops."""
        p = _f64(params)
        n = _f64(norms) if norms is not None else None
        _check(lib().gf_batch_add_point(self.h, t, _dp(p), _dp(n) if n is not None else None))

    def set_norm_ops(self, ops, template, table=()):
        """ops: list of dicts (kind "scale" | "exp" | "dalitz", out, p, n, g, c0, c1); [] removes them."""
        kinds = {"scale": 1, "exp": 2, "dalitz": 3}
        rec = np.zeros(len(ops), dtype=NORM_OP_DTYPE)
        for i, op in enumerate(ops):
            rec[i] = (kinds[op["kind"]], op["out"], op["p"], op.get("n", 0), op.get("g", 0), 0, op.get("c0", 0.0),
                      op.get("c1", 0.0))
        tm = _f64(template)
        tb = _f64(table) if len(table) else np.zeros(1)
        _check(lib().gf_batch_set_norm_ops(self.h, rec.ctypes.data_as(ctypes.c_void_p) if len(ops) else None,
                                           len(ops), _dp(tm), _dp(tb), len(table)))

    def reduce_points(self, metric="NLL"):
        """Evaluate every added point: [ntoys, BATCH_MAX_POINTS] array, NaN where no point was added."""
        out = np.full(self.n * BATCH_MAX_POINTS, np.nan)
        m = METRIC[metric] if isinstance(metric, str) else metric
        _check(lib().gf_batch_reduce_points(self.h, m, _dp(out)))
        return out.reshape(self.n, BATCH_MAX_POINTS)

    def clear_points(self):
        _check(lib().gf_batch_clear_points(self.h))

    def submit_points(self, metric="NLL"):
        t = ctypes.c_uint32()
        m = METRIC[metric] if isinstance(metric, str) else metric
        _check(lib().gf_batch_submit_points(self.h, m, ctypes.byref(t)))
        return t.value

    def wait_points(self, ticket):
        out = np.full(self.n * BATCH_MAX_POINTS, np.nan)
        _check(lib().gf_batch_wait_points(self.h, ticket, _dp(out)))
        return out.reshape(self.n, BATCH_MAX_POINTS)

    def set_chunk(self, chunk):
        _check(lib().gf_batch_set_chunk(self.h, chunk))

    def use_graph(self, on):
        _check(lib().gf_batch_use_graph(self.h, 1 if on else 0))

    def autotune(self, metric="NLL"):
        c, us = ctypes.c_int32(), ctypes.c_double()
        m = METRIC[metric] if isinstance(metric, str) else metric
        _check(lib().gf_batch_autotune(self.h, m, ctypes.byref(c), ctypes.byref(us)))
        return c.value, us.value
