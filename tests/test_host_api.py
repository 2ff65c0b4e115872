"""The GooFit C++ API's host logic on the CPU: data sets, grids, the flattened
tree, parameter handling, and that the device path fails loudly without a GPU.

Known answers from the reference's tests/simple/SimpleTest.cpp and
NormalizeTest.cpp; the flattened arrays are compared with the oracle's
restatement of PdfBase::populateArrays (src/PDFs/PdfBase.cu:268-311)."""
import numpy as np
import pytest

import goofit_amd as G
import oracle as O
from goofit_amd import workloads as W

import _models
from _models import add1d, dalitz as oracle_dalitz, prod3d

FLT_EPS = float(np.finfo(np.float32).eps)


def catch_approx(a, e):
    return abs(a - e) <= 100 * FLT_EPS * abs(e)


def test_add_events_by_value():
    """SimpleTest.cpp:11-45: addEvent() copies the observables' current values."""
    x = G.Observable("xvar", 0, 10)
    y = G.Observable("yvar", 0, 10)
    d = G.UnbinnedDataSet([x, y])
    for a, b in ((1, 2), (3, 4), (5, 6)):
        x.value, y.value = a, b
        d.addEvent()
    assert d.getNumEvents() == 3
    d.loadEvent(1)
    assert x.value == 3 and y.value == 4
    assert [d.getValue(x, i) for i in range(3)] == [1, 3, 5]
    assert [d.getValue(y, i) for i in range(3)] == [2, 4, 6]


def test_make_grid_order():
    """SimpleTest.cpp:81-127: 100 x 100 grid, first observable fastest; EventNumber counts rows."""
    x = G.Observable("xvar", 0, 10)
    y = G.Observable("yvar", 0, 10)
    n = G.EventNumber("eventNumber")
    d = G.UnbinnedDataSet([x, y, n])
    d.fillWithGrid()
    assert d.getNumEvents() == 100 * 100
    for row, (xv, yv) in {0: (0.05, 0.05), 1: (0.15, 0.05), 100: (0.05, 0.15)}.items():
        d.loadEvent(row)
        assert catch_approx(x.value, xv) and catch_approx(y.value, yv)
        assert n.value == float(row)


def test_pdf_make_grid_kat():
    """NormalizeTest.cpp:57-91 and 105-140: makeGrid over each PDF's observables."""
    x = G.Observable("xvar", -5, 10)
    y = G.Observable("yvar", -5, 10)
    g1 = G.GaussianPdf("g1", x, G.Variable("mean1", 1.5), G.Variable("sigma1", 2.5))
    g2 = G.GaussianPdf("g2", y, G.Variable("mean2", -1), G.Variable("sigma2", 0.5))
    d1 = g1.makeGrid()
    for i, v in {0: -4.925, 20: -1.925, 40: 1.075, 60: 4.075}.items():
        assert catch_approx(d1.getValue(x, i), v)
    p = G.ProdPdf("product", [g1, g2])
    d2 = p.makeGrid()
    assert d2.getNumEvents() == 10000
    for i, v in {0: -4.925, 20: -1.925, 40: 1.075, 60: 4.075}.items():
        assert catch_approx(d2.getValue(x, i), v)
        assert catch_approx(d2.getValue(y, 100 * i), v)


def _flat_equal(top_oracle):
    f, p, c, o, n = G.engine.flattened()
    rf, rp, rc, ro, rn = O.flatten(top_oracle).arrays()
    assert list(f) == list(rf)
    assert np.array_equal(np.asarray(p), rp)
    assert np.array_equal(np.asarray(c), rc)
    assert np.array_equal(np.asarray(o), ro)
    assert len(n) == len(rn)


def test_flattened_tree_add1d():
    m = W.build_add1d()
    m.pdf.SigGenSetIndices()
    top, _ = add1d(mu=m.params["mu"].value, sigma=m.params["sigma"].value, alpha=m.params["alpha"].value,
                   frac=m.params["frac"].value)
    _flat_equal(top)


def test_flattened_tree_prod3d():
    m = W.build_prod3d()
    m.pdf.SigGenSetIndices()
    top, _ = prod3d(mu=m.params["mu"].value, sigma=m.params["sigma"].value, alpha=m.params["alpha"].value,
                    rate=m.params["rate"].value)
    _flat_equal(top)


def test_flattened_tree_dalitz():
    """Amp3Body with 16 resonances (Amp3Body.cu:127-196): resonance blocks, efficiency last."""
    m = W.build_dalitz(nbins=200)
    m.pdf.SigGenSetIndices()
    top, _ = oracle_dalitz(nbins=200)
    _flat_equal(top)
    names = G.engine.function_names()
    assert len(names) == 18  # Amp3Body + 16 resonances + efficiency


@pytest.mark.parametrize("kind", ["truth", "threegauss"])
def test_flattened_tree_tddp(kind):
    """Amp3Body_TD (Amp3Body_TD.cu:262-359, 472-529): mixing parameters and couplings,
    constants (nres, cacheToUse, mistag flag, function index after the efficiency),
    resonance blocks, the resolution, then the efficiency."""
    G.Amp3Body_TD.resetCacheCounter()
    m = W.build_tddp(nbins=150, resolution=kind, mistag=True, charmtag=True)
    m.pdf.SigGenSetIndices()
    res = _models.threegauss() if kind == "threegauss" else O.TruthResolution()
    top, _ = _models.tddp(nbins=150, mix=W.TD_MIX, resolution=res, mistag=True, charmtag=True)
    _flat_equal(top)
    f, _, c, _, _ = G.engine.flattened()
    assert c[4] == len(f) and c[3] == 0  # end of the efficiency subtree; mistag observable given
    # the per-D0-mass-bin resolution variant is refused (its device code never selects)
    with pytest.raises(G.GeneralError):
        G.Amp3Body_TD("bad", *m.observables[:4], m.observables[4], m.pdf.getDecayInfo(), [G.TruthResolution()],
                      None, G.Observable("md0", 1.8, 1.9))


def test_minuit_vector_roundtrip_and_blinding():
    """Params::from_minuit_vector (Params.cpp:47-57): values minus the blind offset reach the tree, and
    only changed values are flagged (BlindTest.cpp:48-91 semantics)."""
    m = W.build_add1d()
    fm = G.FitManager(m.pdf)
    par = fm.params
    x = par.make_minuit_vector()
    assert len(x) == par.size()
    x2 = list(x)
    x2[0] += 0.25
    par.from_minuit_vector(x2)
    assert par.make_minuit_vector()[0] == pytest.approx(x2[0])
    f, p, _, _, _ = G.engine.flattened()
    assert x2[0] in list(p)  # host_parameters updated in place (PdfBase::updateParameters)


def test_params_plan_sets_every_slot_and_follows_relayouts():
    """Params::from_minuit_vector replays the tree walks of PdfBase::updateVariable and
    updateParameters (PdfBase.cu:121-140) from a plan: a variable shared by two components
    reaches both parameter slots, blinded values arrive minus the offset, and a new layout
    of the flattened arrays (setIndices, here with a third component in front) is planned
    anew instead of writing the old slots."""
    x = G.Observable("x", -5.0, 5.0)
    mu = G.Variable("mu", 0.1, 0.01, -1.0, 1.0)
    s1 = G.Variable("s1", 1.0, 0.01, 0.1, 3.0)
    s2 = G.Variable("s2", 1.5, 0.01, 0.1, 3.0)
    f = G.Variable("f", 0.4, 0.01, 0.0, 1.0)
    g1, g2 = G.GaussianPdf("g1", x, mu, s1), G.GaussianPdf("g2", x, mu, s2)
    top = G.AddPdf("sum", [f], [g1, g2])
    top.SigGenSetIndices()
    par = G.Params(top)
    names = [v.name for v in top.getParameters()]
    v = list(par.make_minuit_vector())
    v[names.index("mu")] = 0.37
    v[names.index("s1")] = 1.25
    s2.blind = 0.5
    v[names.index("s2")] = 2.0
    par.from_minuit_vector(v)
    assert mu.value == 0.37 and s1.value == 1.25 and s2.value == 1.5
    p = list(G.engine.flattened()[1])
    assert p.count(0.37) == 2 and 1.25 in p and 1.5 in p
    # a new layout: the plan is rebuilt, every slot of the new arrays receives its value
    e0 = G.Variable("e0", -0.2)
    ex = G.ExpPdf("ex", x, e0)
    top2 = G.AddPdf("sum2", [G.Variable("f2", 0.3, 0.01, 0.0, 1.0)], [ex, top])
    top2.SigGenSetIndices()
    v[names.index("mu")] = -0.11
    par.from_minuit_vector(v)
    p2 = list(G.engine.flattened()[1])
    assert p2.count(-0.11) == 2 and 0.37 not in p2
    assert -0.2 in p2  # the new component's own block is intact


def test_device_path_fails_loudly_without_gpu():
    """No CPU fallback: without a GPU every device entry point raises."""
    from goofit_amd import capi
    if capi.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(capi.GooFitError) as e:
        capi.Context(0)
    assert e.value.status == capi.GF_ERR_NODEVICE
    m = W.build_add1d()
    d = G.UnbinnedDataSet(m.observables)
    d.from_columns([W.sample_add1d(100, seed=1)])
    with pytest.raises(Exception):
        m.pdf.setData(d)
        m.pdf.calculateNLL()


def test_binned_dataset_numbering():
    """BinnedDataSet.cpp:37-160: floor((clamped value - lower) / size), first observable fastest,
    centres size * (local + 0.5) + lower, errors sqrt(content) unless set."""
    x = G.Observable("x", 0, 10)
    y = G.Observable("y", -1, 1)
    x.numbins, y.numbins = 10, 4
    d = G.BinnedDataSet(x, y)
    assert d.getNumBins() == 40 and d.getDiminsions() == [10, 4]
    for xv, yv, want in ((0.0, -1.0, 0), (9.999, -1.0, 9), (0.5, -0.49, 10), (3.2, 0.7, 3 + 3 * 10)):
        x.value, y.value = xv, yv
        assert d.getBinNumber() == want
        d.addEvent()
    x.value, y.value = 3.3, 0.9
    d.addWeightedEvent(2.5)
    assert d.getBinContent(33) == 3.5 and d.getNumEvents() == 5
    assert d.getNumWeightedEvents() == 6  # accumulated into an integer, as the reference does
    assert catch_approx(d.getBinCenter(0, 33), 3.5) and catch_approx(d.getBinCenter(1, 33), 0.75)
    assert d.getBinVolume(7) == pytest.approx(1.0 * 0.5)
    assert d.getBinError(33) == pytest.approx(3.5 ** 0.5)
    d.setBinError(33, 0.25)
    assert d.getBinError(33) == 0.25
    # hazard A14: a value equal to the upper limit gets local bin n.  In 2-D that is
    # another bin's number (10 + 2 * 10 = 30, i.e. x-bin 0 of the next y row), as in
    # the reference; in 1-D it is outside the histogram and addEvent throws.
    x.value, y.value = 10.0, 0.0
    assert d.getBinNumber() == 30
    d1 = G.BinnedDataSet(x)
    with pytest.raises(IndexError):
        d1.addEvent()


def test_python_dataset_surface():
    """python/goofit/{UnbinnedDataSet,Variable,DataSet}.cpp of the reference: numpy round trips,
    from_matrix with the event number as the missing last row, properties."""
    x = G.Observable("x", 0, 1)
    y = G.Observable("y", -1, 1)
    n = G.EventNumber("n")
    d = G.UnbinnedDataSet(x, y, n)
    m = np.array([[0.1, 0.2, 1.5, 0.3], [0.0, 0.5, 0.2, -2.0]])
    d.from_matrix(m, filter=True)  # events 3 and 4 are out of range
    assert len(d) == 2 and d.name != "" and [v.name for v in d.variables] == ["x", "y", "n"]
    a = d.to_numpy()
    np.testing.assert_array_equal(a, [[0.1, 0.2], [0.0, 0.5], [0.0, 1.0]])
    np.testing.assert_array_equal(d[1], [0.0, 0.5])
    x.numbins = 4
    assert x.bin_size == 0.25 and x.bin_centers == [0.125, 0.375, 0.625, 0.875]
    v = G.Variable("v", 1.0, 0.1, 0, 2)
    v.blind = 0.5
    assert v.fitterIndex == -1
