"""gf_reduce_metric_noalloc (the reduction for callers that cannot allocate) against
gf_reduce_metric on the same context: the same kernels with the partials in the
library's fixed scratch, so every sum must be the same bits.  Two calls in a row, the
second over a smaller event set, on the context's stream and on a caller's stream; the
coherent-sum Dalitz path and the interpreter path; and the oracle NLL within 1e-11."""
import ctypes
import math
import struct

import numpy as np
import pytest

import oracle as O
from goofit_amd import capi
from goofit_amd import workloads as W

from _models import arrays, dalitz, dalitz_events, prod3d

pytestmark = pytest.mark.gpu

NLL_RTOL = 1e-11


class _Hip:
    """The caller's side: device memory and a stream from the HIP runtime the library
    itself runs on (the process's libamdhip64), through plain C calls."""

    def __init__(self):
        capi.lib()  # the engine's runtime is loaded first
        self.L = ctypes.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipFree.argtypes = [ctypes.c_void_p]
        self.L.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.L.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.L.hipStreamDestroy.argtypes = [ctypes.c_void_p]
        self.L.hipStreamSynchronize.argtypes = [ctypes.c_void_p]

    def check(self, rc, what):
        assert rc == 0, f"{what} failed with HIP error {rc}"

    def malloc(self, nbytes):
        p = ctypes.c_void_p()
        self.check(self.L.hipMalloc(ctypes.byref(p), nbytes), "hipMalloc")
        self.check(self.L.hipMemset(p, 0, nbytes), "hipMemset")
        return p.value

    def free(self, p):
        self.check(self.L.hipFree(ctypes.c_void_p(p)), "hipFree")

    def stream(self):
        s = ctypes.c_void_p()
        self.check(self.L.hipStreamCreate(ctypes.byref(s)), "hipStreamCreate")
        return s.value

    def destroy_stream(self, s):
        self.check(self.L.hipStreamDestroy(ctypes.c_void_p(s)), "hipStreamDestroy")

    def sync(self):
        self.check(self.L.hipDeviceSynchronize(), "hipDeviceSynchronize")

    def read_block(self, dev, stream=None):
        """(sum, residual, error word, sequence number) of a 32-byte device result block."""
        if stream:
            self.check(self.L.hipStreamSynchronize(ctypes.c_void_p(stream)), "hipStreamSynchronize")
        raw = ctypes.create_string_buffer(32)
        self.check(self.L.hipMemcpy(raw, ctypes.c_void_p(dev), 32, 2), "hipMemcpy")  # device to host
        return struct.unpack_from("<ddiI", raw.raw)


@pytest.fixture
def hip():
    return _Hip()


def _sync(ctx, M):
    M.op.sync_params()
    M.op.sync_norms()
    ctx.update_params(M.op.p)
    ctx.update_norms(M.op.n)


@pytest.fixture
def own_ctx():
    c = capi.Context(0)
    yield c
    c.close()


def test_noalloc_dalitz_two_calls(own_ctx, hip):
    ctx = own_ctx
    block = hip.malloc(32)
    stream = hip.stream()
    seqs = []
    for k, n in enumerate((40_000, 9_000)):  # the second call over fewer events
        print(f"noalloc Dalitz: call {k}, {n} events", flush=True)
        top, _ = dalitz(nbins=96)
        aos = dalitz_events(n, 31 + k)
        M = O.Model(top, aos, decay=W.DALITZ_DECAY)
        ref = M.calculate_nll()
        _, (f, p, c, o, nrm) = arrays(top)
        ctx.set_program(f, p, c, o, nrm)
        ctx.set_decay(W.DALITZ_DECAY)
        ctx.upload_columns([aos[:, 0], aos[:, 1], aos[:, 2]])
        ctx.set_identity_column(2)
        ctx.dalitz_set_data_size(top.function_idx, n)
        M.op.sync_params()
        ctx.update_params(M.op.p)
        ctx.dalitz_cache(top.function_idx)
        _sync(ctx, M)
        want, want_c = ctx.reduce_metric("NLL")
        # the first call on the context's own stream, the second on the caller's
        ctx.reduce_metric_noalloc(block, "NLL", stream if k == 1 else 0)
        hip.sync()
        got, got_c, err, seq = hip.read_block(block, stream)
        assert err == 0
        assert got == want and got_c == want_c, (got, want, got_c, want_c)
        assert math.isclose(got, ref, rel_tol=NLL_RTOL), (got, ref)
        seqs.append(seq)
    assert seqs == [1, 2]
    hip.destroy_stream(stream)
    hip.free(block)


def test_noalloc_interpreter_two_calls(own_ctx, hip):
    ctx = own_ctx
    block = hip.malloc(32)
    stream = hip.stream()
    for k, n in enumerate((300_000, 70_001)):
        print(f"noalloc ProdPdf: call {k}, {n} events", flush=True)
        top, obs = prod3d()
        cols = [np.ascontiguousarray(c) for c in W.sample_prod3d(n, seed=50 + k)]
        M = O.Model(top, np.stack(cols, 1))
        ref = M.calculate_nll()
        _, (f, p, c, o, nrm) = arrays(top)
        ctx.set_program(f, p, c, o, nrm)
        ctx.upload_columns(cols)
        _sync(ctx, M)
        want, want_c = ctx.reduce_metric("NLL")
        ctx.reduce_metric_noalloc(block, "NLL", stream if k == 0 else 0)
        hip.sync()
        got, got_c, err, _ = hip.read_block(block, stream)
        assert err == 0
        assert got == want and got_c == want_c, (got, want, got_c, want_c)
        assert math.isclose(got, ref, rel_tol=NLL_RTOL), (got, ref)
    hip.destroy_stream(stream)
    hip.free(block)


def test_noalloc_refuses_a_program_too_large_for_the_kernel_arguments(own_ctx, hip):
    """A parameter array beyond the kernel-argument block would need the staging copy
    (an allocation): the call is refused with a clear error."""
    ctx = own_ctx
    ctx.set_tuning(capi.TUNE_INLINE_ARGS, 0)
    try:
        top, obs = prod3d()
        cols = [np.ascontiguousarray(c) for c in W.sample_prod3d(1000, seed=5)]
        M = O.Model(top, np.stack(cols, 1))
        _, (f, p, c, o, nrm) = arrays(top)
        ctx.set_program(f, p, c, o, nrm)
        ctx.upload_columns(cols)
        _sync(ctx, M)
        block = hip.malloc(32)
        with pytest.raises(capi.GooFitError, match="kernel"):
            ctx.reduce_metric_noalloc(block, "NLL")
        hip.free(block)
    finally:
        ctx.set_tuning(capi.TUNE_INLINE_ARGS, 1)
