"""The drop-in estimator classes on the device against the oracle, with the
reference's Python call pattern (python/boitata/estimation.py:405-485):
fit / predict / cross_validate, search passes and shared locations."""
import warnings

import numpy as np
import pytest

import oracle
from boitata_amd import _abi, datasets
from boitata_amd.estimation import (BlockKriging, IndicatorKriging, OrdinaryKriging, Search, SimpleKriging,
                                    Variogram)

pytestmark = pytest.mark.gpu


def close(a, b, tol, floor):
    a, b = np.asarray(a), np.asarray(b)
    assert (np.isnan(a) == np.isnan(b)).all()
    ok = ~np.isnan(b)
    err = np.abs(a[ok] - b[ok]) / np.maximum(np.abs(b[ok]), floor)
    assert err.max(initial=0.0) <= tol, err.max()


@pytest.fixture(scope="module")
def c1():
    w = datasets.config1()
    P = np.stack([w.x, w.y, w.z], 1)
    return w, P, w.targets[::41].copy()


def vg_c1():
    return Variogram([("spherical", 0.8, 120)], nugget=0.2, rotation=(30, 0, 0), ratios=(80 / 120, 30 / 120))


def oracle_passes(d, vg, searches, T, fn):
    value = np.full(len(T), np.nan)
    var = np.full(len(T), np.nan)
    rem = np.arange(len(T))
    for s in searches:
        if len(rem) == 0:
            break
        o = fn(d, vg, s, T[rem])
        ok = o["status"] == 0
        value[rem[ok]] = o["value"][ok]
        var[rem[ok]] = o["variance"][ok]
        rem = rem[~ok]
    return value, var


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("fp32_ir", 1e-6)])
def test_ordinary_kriging_predict(c1, precision, tol):
    w, P, T = c1
    vg = vg_c1()
    s = Search(ranges=(120, 80, 30), rotation=(30, 0, 0), min_samples=4, max_samples=24)
    est = OrdinaryKriging(vg, s, precision=precision).fit(P, w.value, holes=w.hole)
    v, q = est.predict(T, return_variance=True)
    o = oracle.krige(oracle.Data(w.x, w.y, w.z, w.value, None, w.hole), vg._c, s._c, T)
    close(v, np.where(o["status"] == 0, o["value"], np.nan), tol, 1e-3 * vg.sill)
    close(q, np.where(o["status"] == 0, o["variance"], np.nan), tol, 1e-3 * vg.sill)
    assert np.array_equal(est.predict(T), v, equal_nan=True)


def test_passes_and_simple_and_indicator(c1):
    w, P, T = c1
    vg = vg_c1()
    strict = Search(60.0, min_samples=20, max_samples=24)
    wide = Search(ranges=(150, 100, 40), rotation=(30, 0, 0), min_samples=2, max_samples=16)
    d = oracle.Data(w.x, w.y, w.z, w.value, None, w.hole)
    for cls, kw, kind, param in ((OrdinaryKriging, {}, _abi.ORDINARY, 0.0),
                                 (SimpleKriging, dict(mean=1.0), _abi.SIMPLE, 1.0),
                                 (IndicatorKriging, dict(threshold=1.2), _abi.INDICATOR, 1.2)):
        args = (vg, [strict, wide]) if cls is not IndicatorKriging else (vg, [strict, wide], 1.2)
        kw = kw if cls is not IndicatorKriging else {}
        est = cls(*args, **kw).fit(P, w.value, holes=w.hole)
        v, q = est.predict(T, return_variance=True)
        ov, oq = oracle_passes(d, vg._c, [strict._c, wide._c], T,
                               lambda d_, g, s, t: oracle.krige(d_, g, s, t, kind, param))
        close(v, ov, 1e-10, 1e-3 * vg.sill)
        close(q, oq, 1e-10, 1e-3 * vg.sill)
        first = oracle.krige(d, vg._c, strict._c, T)["status"]
        assert (first != 0).any() and (first == 0).any()  # both passes did work


def test_cross_validate_is_leave_one_out(c1):
    w, P, _ = c1
    vg = vg_c1()
    s = Search(ranges=(120, 80, 30), rotation=(30, 0, 0), min_samples=4, max_samples=24)
    est = OrdinaryKriging(vg, s).fit(P, w.value, holes=w.hole)
    cv = est.cross_validate()
    o = oracle.loo(oracle.Data(w.x, w.y, w.z, w.value, None, w.hole), vg._c, s._c, np.arange(w.n))
    close(cv.estimate, np.where(o["status"] == 0, o["value"], np.nan), 1e-10, 1e-3 * vg.sill)
    assert np.array_equal(cv.actual, w.value)
    assert 0.0 < cv.standardized_squared_error < 10.0 and cv.correlation > 0.3


def test_block_kriging(c1):
    w, P, T = c1
    vg = vg_c1()
    s = Search(ranges=(120, 80, 30), rotation=(30, 0, 0), min_samples=4, max_samples=24)
    est = BlockKriging(vg, s, (10, 10, 10), discretization=(4, 4, 2)).fit(P, w.value, holes=w.hole)
    v, q = est.predict(T, return_variance=True)
    o = oracle.block_krige(oracle.Data(w.x, w.y, w.z, w.value, None, w.hole), vg._c, s._c,
                           _abi.make_block((10, 10, 10), (4, 4, 2)), T)
    close(v, np.where(o["status"] == 0, o["value"], np.nan), 1e-10, 1e-3 * vg.sill)
    close(q, np.where(o["status"] == 0, o["variance"], np.nan), 1e-10, 1e-3 * vg.sill)


def test_shared_locations_are_dropped_before_the_device():
    rng = np.random.default_rng(3)
    P = rng.uniform(0, 100, (400, 3))
    v = rng.uniform(0, 2, 400)
    P2 = np.concatenate([P, P[:5]])
    v2 = np.concatenate([v, v[:5] + 1.0])
    vg = Variogram([("exponential", 1.0, 40.0)], nugget=0.1)
    s = Search(60.0, max_samples=20)
    T = rng.uniform(0, 100, (200, 3))
    with warnings.catch_warnings(record=True) as wr:
        warnings.simplefilter("always")
        a = OrdinaryKriging(vg, s).fit(P2, v2).predict(T)
    assert any("rows 0, 400" in str(x.message) for x in wr)
    b = OrdinaryKriging(vg, s).fit(P, v).predict(T)
    assert np.array_equal(a, b, equal_nan=True)
