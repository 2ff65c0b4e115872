"""goofit_amd -- GooFit's likelihood hot path on AMD Instinct MI355X (gfx950).

The GooFit C++ API (GooPdf, AddPdf, ProdPdf, Amp3Body, Amp3Body_TD, UnbinnedDataSet,
FitManager, ...) is built as libgoofit.so on top of the C ABI of
include/goofit_hip.h (libgoofit_hip.so, hand-written HIP kernels).  This package
exposes that API to Python (the `_goofit` binding) and the raw C ABI (`capi`).

Importing the API requires the native libraries to have been built in-tree
(`python -m goofit_amd.build`); there is no CPU fallback.
"""
from __future__ import annotations

import os

_HERE = os.path.dirname(os.path.abspath(__file__))

try:
    from . import _goofit  # noqa: F401
    from ._goofit import (  # noqa: F401
        AddPdf, Amp3Body, Amp3Body_TD, BatchFitter, BinnedChisqFit, BinnedDataSet, BinnedErrorFit, BinnedNllFit, DalitzProposal,
        DataSet, EvalFit, FitControl, ProbFit, UnbinnedNllFit, DaughterPair, DecayInfo3, DecayInfo3t, EvalFunc, EventNumber, ExpGausPdf, ExpPdf,
        FCN, FitManager, FunctionMinimum, GaussianPdf, GeneralError, GooPdf, Observable, PAIR_12, PAIR_13, PAIR_23,
        MixingTimeResolution, Params, PdfBase, PolynomialPdf, ProdPdf, Resonances, ThreeGaussResolution,
        TruthResolution, UnbinnedDataSet, Variable, WaveHolder_s, engine, inDalitz,
    )
    HAVE_API = True
except ImportError as _e:  # the binding is not built yet
    HAVE_API = False
    _IMPORT_ERROR = _e


def require_api():
    """Raise if the compiled API is missing (the product never falls back to Python)."""
    if not HAVE_API:
        raise ImportError(f"goofit_amd native API not built ({_IMPORT_ERROR}); run `python -m goofit_amd.build`")


def native_libraries():
    """Paths of the in-tree native libraries this package loads."""
    lib = os.path.join(_HERE, "lib")
    return [os.path.join(lib, "libgoofit_hip.so"), os.path.join(lib, "libgoofit.so")]
