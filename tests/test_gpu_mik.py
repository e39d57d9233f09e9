"""Multiple indicator kriging on the device against the oracle
(crates/estimation/src/indicator.rs:439-529): conditional distributions of
the reference's own test inputs (LCG clouds, six thresholds with their own
variograms, every tail and interpolation model), simple kriging with
declustering weights, block support, leave-one-out with the PIT, and a
drill-hole workload with decile thresholds."""
import numpy as np
import pytest

import oracle
from boitata_amd import _abi, datasets
from test_oracle_mik import data, grid, mik, search, single

pytestmark = pytest.mark.gpu

KEYS = ("raw", "cdf", "correction", "mean", "variance", "probability_above", "mean_above", "quantile_values")


def assert_same(g, o, tol, floor=1e-3, keys=KEYS):
    """Device against oracle. mean_above is NaN where P(Z > cutoff) is 0; when
    every neighbour lies below the top threshold that probability is 1 minus a
    sum of weights equal to 1 up to rounding, so it is compared only where the
    probability is not within rounding of 0 (each side NaN exactly where its
    own probability is 0)."""
    assert np.array_equal(g["status"], o["status"])
    ok = o["status"] == 0
    assert ok.any()
    assert np.array_equal(g["n_used"][ok], o["n_used"][ok])
    for k in keys:
        a, b = np.asarray(g[k]), np.asarray(o[k])
        a, b = a[..., ok], b[..., ok]
        if k == "mean_above":
            pa, pb = np.asarray(g["probability_above"])[..., ok], np.asarray(o["probability_above"])[..., ok]
            assert np.array_equal(np.isnan(a), pa <= 0.0) and np.array_equal(np.isnan(b), pb <= 0.0)
            firm = (pa > 1e-12) & (pb > 1e-12)
            a, b = a[firm], b[firm]
        assert (np.isnan(a) == np.isnan(b)).all(), k
        fin = ~np.isnan(b)
        err = np.abs(a[fin] - b[fin]) / np.maximum(np.abs(b[fin]), floor)
        assert err.max(initial=0.0) <= tol, (k, err.max())
    for k in KEYS:
        if k in g and np.asarray(g[k]).size:
            assert np.isnan(np.asarray(g[k])[..., ~ok]).all(), k


def engine_with(engine, d):
    engine.set_samples(d.x, d.y, d.z, d.value, None, d.hole)
    return engine


@pytest.mark.parametrize("precision,tol", [(_abi.FP64, 1e-10), (_abi.FP32_IR, 1e-6)])
@pytest.mark.parametrize("tail,interp", [(_abi.TAIL_NONE, _abi.MIK_GLOBAL), (_abi.TAIL_NONE, _abi.MIK_LINEAR),
                                         ((_abi.TAIL_POWER, 0.5), _abi.MIK_GLOBAL),
                                         ((_abi.TAIL_HYPERBOLIC, 1.5), _abi.MIK_LINEAR)])
def test_distributions_match_the_oracle(engine, tail, interp, precision, tol):
    # Six thresholds with their own variograms, nugget 0.1 of a sill of 1 (the
    # systems stay well conditioned, so FP64 parity holds at 1e-10).
    d = data(150, 7)
    thr = [0.2, 0.5, 0.8, 1.2, 1.8, 2.6]
    models = ("gaussian", "spherical", "exponential", "gaussian", "spherical", "exponential")
    vgs = [_abi.make_variogram([(mo, 0.9, r)], nugget=0.1) for mo, r in zip(models, (10.0, 60.0, 15.0, 80.0, 5.0,
                                                                                         40.0))]
    kind, w = (tail, 1.0) if tail == _abi.TAIL_NONE else tail
    p = mik(thr, 6, interpolation=interp, upper_tail=kind, omega=w)
    # Interior quantiles: at 0 and 1 the class chosen depends on whether a sum
    # of weights rounds to exactly 1 (the oracle tests cover the end points).
    kw = dict(cutoffs=[0.5, 1.0, 3.0], quantiles=[0.1, 0.5, 0.9])
    T = grid()
    o = oracle.mik_predict(d, p, thr, vgs, search(40.0), T, **kw)
    g = engine_with(engine, d).mik_predict(p, thr, vgs, search(40.0), T[:, 0], T[:, 1], T[:, 2],
                                           precision=precision, **kw)
    assert_same(g, o, tol)
    # Violations count entries the correction changed (cdf != raw). A kriged
    # probability equal to 0 or 1 up to rounding is clipped, or swapped with a
    # neighbour, by one ulp on one side and not on the other, so the counts
    # are compared over the corrections larger than 1e-12.
    def firm(r):
        return (np.abs(r["cdf"] - r["raw"]) > 1e-12).sum(axis=0)
    assert np.array_equal(firm(g), firm(o))
    assert np.all(g["violations"] >= firm(g)) and np.all(o["violations"] >= firm(o))
    assert (o["correction"] > 0).any()


def gaussian_condition(d, vgs_ranges, search_vg, T, radius):
    """Largest 2-norm condition number, over each target's variograms, of its
    kriging matrix (unit-sill Gaussian, no nugget) on the neighbours the oracle
    selects."""
    idx, cnt, _ = oracle.neighbors(d, search(radius), search_vg, T)
    P = np.c_[d.x, d.y, d.z]
    kappa = np.ones(len(T))
    for i in range(len(T)):
        q = P[idx[i, : cnt[i]]]
        h = np.sqrt(((q[:, None, :] - q[None, :, :]) ** 2).sum(-1))
        for r in vgs_ranges:
            C = np.exp(-3.0 * (h / r) ** 2)
            kappa[i] = max(kappa[i], np.linalg.cond(C))
    return kappa


@pytest.mark.parametrize("precision", [_abi.FP64, _abi.FP32_IR])
def test_reference_inputs_within_their_conditioning(engine, precision):
    # indicator.rs:933-982 verbatim: unit-sill Gaussian variograms without a
    # nugget. Their systems are ill-conditioned (condition numbers up to ~1e9
    # here), so a last-bit difference between the device's and the host's exp
    # grows by the condition number: the bound is 1e-10 + 1e-14 kappa per target
    # (1e-6 + 1e-14 kappa for FP32 + refinement) on the kriged and corrected
    # probabilities and the moments.
    d = data(150, 7)
    thr = [0.2, 0.5, 0.8, 1.2, 1.8, 2.6]
    ranges = (10.0, 60.0, 15.0, 80.0, 5.0, 40.0)
    vgs = [single("gaussian", 1.0, r) for r in ranges]
    p = mik(thr, 6)
    T = grid()
    o = oracle.mik_predict(d, p, thr, vgs, search(40.0), T, cutoffs=[1.0])
    g = engine_with(engine, d).mik_predict(p, thr, vgs, search(40.0), T[:, 0], T[:, 1], T[:, 2], cutoffs=[1.0],
                                           precision=precision)
    assert np.array_equal(g["status"], o["status"]) and (o["status"] == 0).all()
    kappa = gaussian_condition(d, ranges, vgs[3], T, 40.0)
    base = 1e-10 if precision == _abi.FP64 else 1e-6
    bound = base + 1e-14 * kappa
    for k in ("raw", "cdf", "mean", "variance", "probability_above"):
        a, b = np.atleast_2d(g[k]), np.atleast_2d(o[k])
        err = np.abs(a - b) / np.maximum(np.abs(b), 1e-3)
        assert np.all(err <= bound), (k, float((err / bound).max()), float(kappa.max()))


@pytest.mark.parametrize("precision,tol", [(_abi.FP64, 1e-10), (_abi.FP32_IR, 1e-6)])
def test_simple_form_with_declustering_weights(engine, precision, tol):
    d = data(120, 11)
    w = 1.0 + (np.arange(120) % 5)
    thr = [0.3, 0.7, 1.5, 2.5]
    p = mik(thr, 1, simple=True)
    vg = [single("spherical", 1.0, 20.0)]
    T = np.r_[grid()[::3], [[1e4, 1e4, 0.0]]]
    o = oracle.mik_predict(d, p, thr, vg, search(30.0), T, weights=w, cutoffs=[1.0], quantiles=[0.5])
    g = engine_with(engine, d).mik_predict(p, thr, vg, search(30.0), T[:, 0], T[:, 1], T[:, 2], weights=w,
                                           cutoffs=[1.0], quantiles=[0.5], precision=precision)
    assert_same(g, o, tol)
    # Far from the data (one sample in reach of the infinite search) the simple
    # form gives the declustered mean (indicator.rs:1118-1150).
    far = engine.mik_predict(p, thr, vg, search(1e6), [1e4], [1e4], [0.0], weights=w)
    assert abs(far["mean"][0] - (d.value * w).sum() / w.sum()) < 1e-10


def test_one_threshold_matches_indicator_kriging(engine):
    d = data(80, 3)
    vg = single("spherical", 0.25, 30.0)
    T = grid()
    g = engine_with(engine, d).mik_predict(mik([1.0], 1), [1.0], [vg], search(25.0), T[:, 0], T[:, 1], T[:, 2])
    ik = engine.krige(vg, search(25.0), T[:, 0], T[:, 1], T[:, 2], kind=_abi.INDICATOR, param=1.0)
    ok = ik["status"] == 0
    assert np.array_equal(g["status"] == 0, ok) and ok.any()
    assert np.abs(g["cdf"][0, ok] - ik["value"][ok]).max() <= 1e-12


@pytest.mark.parametrize("simple", [False, True])
def test_block_support(engine, simple):
    d = data(60, 11)
    thr = [0.5, 1.0, 2.0]
    vgs = [single("spherical", 0.25, r) for r in (20.0, 40.0, 15.0)]
    p = mik(thr, 3, simple=simple)
    blk = _abi.make_block((10.0, 6.0, 1.0), (4, 4, 1), (30.0, 0.0, 0.0))
    T = grid()[::7]
    o = oracle.mik_predict(d, p, thr, vgs, search(30.0), T, block=blk, cutoffs=[1.0], quantiles=[0.25, 0.75])
    g = engine_with(engine, d).mik_predict(p, thr, vgs, search(30.0), T[:, 0], T[:, 1], T[:, 2], block=blk,
                                           cutoffs=[1.0], quantiles=[0.25, 0.75])
    assert_same(g, o, 1e-9)


@pytest.mark.parametrize("precision,tol", [(_abi.FP64, 1e-10), (_abi.FP32_IR, 1e-6)])
def test_cross_validation_and_pit(engine, precision, tol):
    d = data(150, 7)
    thr = [0.2, 0.5, 0.8, 1.2, 1.8, 2.6]
    vg = [single("exponential", 1.0, 30.0)]
    p = mik(thr, 1)
    o = oracle.mik_cross_validate(d, p, thr, vg, search(25.0))
    g = engine_with(engine, d).mik_cross_validate(p, thr, vg, search(25.0), precision=precision)
    assert_same(g, o, tol, keys=("raw", "cdf", "correction", "mean", "variance", "pit"))
    which = np.arange(0, 150, 4)
    o2 = oracle.mik_cross_validate(d, p, thr, vg, search(25.0), which=which)
    g2 = engine.mik_cross_validate(p, thr, vg, search(25.0), which=which)
    assert_same(g2, o2, 1e-10, keys=("cdf", "mean", "pit"))


def test_drill_hole_deciles(engine):
    # The 5,000-composite drill-hole workload with decile thresholds, one
    # variogram: 50,000 block centroids on the device, every 37th against the oracle.
    w = datasets.config1()
    d = oracle.Data(w.x, w.y, w.z, w.value, None, w.hole)
    thr = list(np.quantile(w.value, np.linspace(0.1, 0.9, 9)))
    p = _abi.make_mik(thr, 1)
    kw = dict(cutoffs=[0.5, 1.0, 2.0], quantiles=[0.1, 0.5, 0.9])
    T = w.targets
    engine.set_samples(w.x, w.y, w.z, w.value, None, w.hole)
    g = engine.mik_predict(p, thr, [w.variogram], w.search, T[:, 0], T[:, 1], T[:, 2], **kw)
    sub = np.arange(0, len(T), 37)
    o = oracle.mik_predict(d, p, thr, [w.variogram], w.search, T[sub], **kw)
    gs = {k: (v[..., sub] if isinstance(v, np.ndarray) else v) for k, v in g.items()}
    assert_same(gs, o, 1e-10)
    ok = g["status"] == 0
    assert ok.mean() > 0.9
    cdf = g["cdf"][:, ok]
    assert np.all((cdf >= 0) & (cdf <= 1)) and np.all(np.diff(cdf, axis=0) >= 0)


def test_device_arrays_and_invalid_parameters(engine):
    torch = pytest.importorskip("torch")
    d = data(150, 7)
    thr = [0.5, 1.0, 2.0]
    vg = [single("exponential", 1.0, 30.0)]
    p = mik(thr, 1)
    T = grid()
    host = engine_with(engine, d).mik_predict(p, thr, vg, search(25.0), T[:, 0], T[:, 1], T[:, 2], cutoffs=[1.0])
    Td = torch.tensor(T, device="cuda:0")
    dev = engine.mik_predict(p, thr, vg, search(25.0), Td[:, 0].contiguous(), Td[:, 1].contiguous(),
                             Td[:, 2].contiguous(), cutoffs=[1.0])
    for k in ("cdf", "mean", "variance", "probability_above", "status"):
        assert np.array_equal(dev[k].cpu().numpy(), host[k], equal_nan=True), k
    from boitata_amd.engine import EngineError

    bad = [([1.0, 1.0], 1, {}), ([0.0, 1.0, 2.0], 2, {}), ([-1.0], 1, dict(upper_tail=_abi.TAIL_HYPERBOLIC,
                                                                             omega=2.0))]
    for t, nv, kw in bad:
        with pytest.raises(EngineError):
            engine.mik_predict(mik(t, nv, **kw), t, vg * nv, search(25.0), T[:3, 0], T[:3, 1], T[:3, 2])
    with pytest.raises(EngineError):
        engine.mik_predict(p, thr, vg, search(25.0), T[:3, 0], T[:3, 1], T[:3, 2], quantiles=[1.5])
    with pytest.raises(EngineError):
        engine.mik_predict(p, thr, vg, _abi.make_search(1, 12, 25.0, target_efficiency=0.5), T[:3, 0], T[:3, 1],
                           T[:3, 2])


def test_estimator_passes_diagnostics_and_cross_validation():
    # indicator.rs:1097-1116 and python/tests/test_estimation.py:575-618: two
    # passes, diagnostics that follow the search and the correction, LOO.
    from boitata_amd.estimation import MultipleIndicatorKriging, Search, Variogram

    d = data(150, 7)
    P = np.c_[d.x, d.y]
    thr = [0.2, 0.5, 0.8, 1.2, 1.8, 2.6]
    vgs = [Variogram([("gaussian", 0.9, r)], nugget=0.1) for r in (10.0, 60.0, 15.0, 80.0, 5.0, 40.0)]
    passes = [Search(6.0, max_samples=12, min_samples=1), Search(40.0, max_samples=12, min_samples=1)]
    est = MultipleIndicatorKriging(vgs, passes, thr, tails=(0.0, 10.0)).fit(P, d.value)
    T = grid()
    s = est.predict(T[:, :2], cutoffs=[1.0], quantiles=[0.1, 0.5, 0.9], diagnostics=True)
    assert s.cdf.shape == (400, 6) and s.quantile_values.shape == (400, 3) and s.probability_above.shape == (400, 1)
    # The passes composed from the oracle.
    p = mik(thr, 6)
    want = np.full(400, np.nan)
    rem = np.arange(400)
    for sp in passes:
        o = oracle.mik_predict(d, p, thr, [v._c for v in vgs], sp._c, T[rem])
        ok = o["status"] == 0
        want[rem[ok]] = o["mean"][ok]
        rem = rem[~ok]
    assert np.array_equal(np.isnan(s.mean), np.isnan(want))
    fin = ~np.isnan(want)
    assert np.abs(s.mean[fin] - want[fin]).max() <= 1e-10 * max(1.0, np.abs(want[fin]).max())
    dg = s.diagnostics
    assert dg.column_names == ["n_samples", "pass", "n_holes", "mean_distance", "max_samples_reached", "correction",
                               "n_order_violations"]
    assert np.array_equal(dg["n_order_violations"] > 0, s.correction > 0)
    assert np.array_equal(dg["max_samples_reached"] == 1.0, dg["n_samples"] == 12.0)
    assert np.all((dg["n_samples"] >= 1) & (dg["n_samples"] <= 12))
    assert set(np.unique(dg["pass"])) == {1.0, 2.0}
    assert np.array_equal(dg["n_holes"], dg["n_samples"])  # no holes: every sample its own
    lim = np.where(dg["pass"] == 1, 6.0, 40.0)
    assert np.all(dg["mean_distance"] <= lim + 1e-12) and np.all(dg["mean_distance"] >= 0)
    assert est.predict(T[:5, :2]).diagnostics is None
    cv = est.cross_validate()
    assert cv.cdf.shape == (150, 6)
    ok = ~np.isnan(cv.pit)
    assert ok.all() and np.all((cv.pit >= 0) & (cv.pit <= 1))
    assert cv.brier.shape == (6,) and np.all((cv.brier >= 0) & (cv.brier <= 1)) and 0.0 < cv.goodness <= 1.0


def test_no_targets(engine):
    d = data(50, 2)
    thr = [0.5, 1.0]
    vg = [single("spherical", 1.0, 30.0)]
    e = engine_with(engine, d)
    out = e.mik_predict(mik(thr, 1), thr, vg, search(25.0), [], [], [], cutoffs=[1.0], quantiles=[0.5])
    assert out["cdf"].shape == (2, 0) and out["status"].shape == (0,)
    cv = e.mik_cross_validate(mik(thr, 1), thr, vg, search(25.0), which=np.zeros(0, np.int64))
    assert cv["pit"].shape == (0,)
