"""Drop-in pieces of the estimation path against the oracle, through the C ABI:
- weights_many (crates/estimation/src/batch.rs:76-95): the samples each target
  is kriged from, in the estimator's order, and their weights;
- leave-one-out block kriging (BlockKriging.cross_validate,
  crates/py/src/estimation.rs:1826-1834 with leave_one_out_at);
- k-fold cross-validation (k_fold_at, batch.rs:197-243), holes kept whole.
Sample lists bit for bit; values, variances and weights at 1e-10 (FP64) and
1e-6 (FP32+IR), relative with a floor of 1e-3 * sill."""
import numpy as np
import pytest

import oracle
from boitata_amd import _abi, datasets

pytestmark = pytest.mark.gpu


def sill_of(vg):
    return vg.nugget + sum(vg.structures[k].sill for k in range(vg.n_structures))


def rel(a, b, floor):
    return np.abs(np.asarray(a) - np.asarray(b)) / np.maximum(np.abs(np.asarray(b)), floor)


@pytest.fixture(scope="module")
def c1():
    w = datasets.config1()
    return w, oracle.Data(w.x, w.y, w.z, w.value, None, w.hole)


@pytest.mark.parametrize("precision,tol", [(_abi.FP64, 1e-10), (_abi.FP32_IR, 1e-6)])
@pytest.mark.parametrize("kind,param", [(_abi.ORDINARY, 0.0), (_abi.SIMPLE, 1.0)])
def test_weights_many(engine, c1, precision, tol, kind, param):
    w, d = c1
    engine.set_samples(w.x, w.y, w.z, w.value, None, w.hole)
    T = w.targets[::97][:300]
    g = engine.krige_weights(w.variogram, w.search, T[:, 0], T[:, 1], T[:, 2], kind, param, precision=precision)
    oi, oc, os_ = oracle.neighbors(d, w.search, w.variogram, T)
    assert (g["status"] == os_).all()
    ok = os_ == 0
    assert (g["count"][ok] == oc[ok]).all() and (g["count"][~ok] == 0).all()
    for t in np.nonzero(ok)[0][::3]:
        n = oc[t]
        assert (g["idx"][t, :n] == oi[t, :n]).all()
        sub = oracle.Data(w.x[oi[t, :n]], w.y[oi[t, :n]], w.z[oi[t, :n]], w.value[oi[t, :n]])
        e = oracle.krige_system(sub, w.variogram, T[t], kind, param)
        assert e["status"] == 0
        floor = 1e-3 * max(np.abs(e["weights"]).max(), 1e-12)
        assert rel(g["weights"][t, :n], e["weights"], floor).max() <= tol, t
    # The weights reproduce the estimate (krige.rs:98-113).
    k = engine.krige(w.variogram, w.search, T[:, 0], T[:, 1], T[:, 2], kind, param, precision)
    for t in np.nonzero(ok)[0]:
        n = oc[t]
        z = w.value[g["idx"][t, :n]]
        est = g["weights"][t, :n] @ (z - param) + param if kind == _abi.SIMPLE else g["weights"][t, :n] @ z
        assert abs(est - k["value"][t]) <= 1e-9 * max(abs(k["value"][t]), 1e-3)


def test_weights_many_blocks_and_coincident_samples(engine):
    rng = np.random.default_rng(11)
    P = rng.uniform(0, 100, (400, 3))
    P[5] = P[6]  # an exact twin: merged, its weight shared (coincident.rs:102-112)
    v = rng.uniform(0, 3, 400)
    engine.set_samples(P[:, 0], P[:, 1], P[:, 2], v)
    d = oracle.Data(P[:, 0], P[:, 1], P[:, 2], v)
    vg = _abi.make_variogram([("spherical", 0.9, 50.0)], nugget=0.1)
    s = _abi.make_search(2, 16, 40.0)
    T = np.concatenate([P[[5, 6]] + 0.3, rng.uniform(10, 90, (60, 3))])
    blk = _abi.make_block((6, 6, 6), (2, 2, 2))
    g = engine.krige_weights(vg, s, T[:, 0], T[:, 1], T[:, 2], block=blk)
    oi, oc, os_ = oracle.neighbors(d, s, vg, T)
    for t in np.nonzero(os_ == 0)[0]:
        n = oc[t]
        sub = oracle.Data(P[oi[t, :n], 0], P[oi[t, :n], 1], P[oi[t, :n], 2], v[oi[t, :n]])
        e = oracle.block_krige_system(sub, vg, oracle.block_points(T[t], blk))
        assert e["status"] == 0
        assert np.allclose(g["weights"][t, :n], e["weights"], rtol=1e-9, atol=1e-12), t


@pytest.mark.parametrize("precision,tol", [(_abi.FP64, 1e-10), (_abi.FP32_IR, 1e-6)])
def test_block_leave_one_out(engine, c1, precision, tol):
    w, d = c1
    engine.set_samples(w.x, w.y, w.z, w.value, None, w.hole)
    blk = _abi.make_block((10, 10, 4), (3, 3, 2), rotation=(20, 0, 0))
    which = np.arange(0, w.n, 11)
    g = engine.block_loo(w.variogram, w.search, blk, which, precision=precision)
    P = np.stack([w.x, w.y, w.z], 1)[which]
    oi, oc, os_ = oracle.neighbors(d, w.search, w.variogram, P, left_out=which)
    o = oracle.block_krige_lists(d, w.variogram, blk, P, oi, oc, solver=0)
    o["status"] = np.where(os_ != 0, os_, np.where(oc == 0, _abi.SEARCH_FAILED, o["status"]))
    assert (g["status"] == o["status"]).all()
    ok = o["status"] == 0
    floor = 1e-3 * sill_of(w.variogram)
    for k in ("value", "variance"):
        assert rel(g[k][ok], o[k][ok], floor).max() <= tol, k


def kfold_oracle(w, vg, search, k, which):
    """k_fold_at (batch.rs:197-243) over the oracle: folds by sorted hole code."""
    codes = np.unique(w.hole)
    fold = np.searchsorted(codes, w.hole) % k
    value, variance = np.full(len(which), np.nan), np.full(len(which), np.nan)
    status = np.zeros(len(which), np.int32)
    P = np.stack([w.x, w.y, w.z], 1)
    for f in range(k):
        test = np.nonzero(fold[which] == f)[0]
        train = fold != f
        dtr = oracle.Data(w.x[train], w.y[train], w.z[train], w.value[train], None, w.hole[train])
        o = oracle.krige(dtr, vg, search, P[which[test]])
        value[test], variance[test], status[test] = o["value"], o["variance"], o["status"]
    return value, variance, status


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("fp32_ir", 1e-6)])
def test_k_fold_cross_validation(c1, precision, tol):
    from boitata_amd.estimation import OrdinaryKriging, Search, Variogram

    w, _ = c1
    vg = Variogram([("spherical", 0.8, 120)], nugget=0.2, rotation=(30, 0, 0), ratios=(80 / 120, 30 / 120))
    est = OrdinaryKriging(vg, Search(ranges=(120, 80, 30), rotation=(30, 0, 0), min_samples=4, max_samples=24),
                          precision=precision)
    est.fit(np.stack([w.x, w.y, w.z], 1), w.value, holes=w.hole)
    cv = est.cross_validate(folds=5)
    ov, oq, ost = kfold_oracle(w, vg._c, est._passes[0]._c, 5, np.arange(w.n))
    ok = ost == 0
    assert (np.isnan(cv.estimate) == ~ok).all()
    floor = 1e-3 * sill_of(vg._c)
    assert rel(cv.estimate[ok], ov[ok], floor).max() <= tol
    assert rel(cv.variance[ok], oq[ok], floor).max() <= tol
    from boitata_amd.errors import InvalidInput

    with pytest.raises(InvalidInput):
        est.cross_validate(folds=1)


def test_block_cross_validate_drop_in(c1):
    from boitata_amd.estimation import BlockKriging, Search, Variogram

    w, d = c1
    vg = Variogram([("spherical", 0.8, 120)], nugget=0.2, rotation=(30, 0, 0), ratios=(80 / 120, 30 / 120))
    s = Search(ranges=(120, 80, 30), rotation=(30, 0, 0), min_samples=4, max_samples=24)
    est = BlockKriging(vg, s, (10, 10, 5), discretization=(2, 2, 2)).fit(np.stack([w.x, w.y, w.z], 1), w.value,
                                                                          holes=w.hole)
    cv = est.cross_validate()
    P = np.stack([w.x, w.y, w.z], 1)
    which = np.arange(w.n)
    oi, oc, os_ = oracle.neighbors(d, s._c, vg._c, P, left_out=which)
    o = oracle.block_krige_lists(d, vg._c, est._block, P, oi, oc, solver=0)
    ok = (os_ == 0) & (oc > 0) & (o["status"] == 0)
    assert (np.isnan(cv.estimate) == ~ok).all()
    assert rel(cv.estimate[ok], o["value"][ok], 1e-3).max() <= 1e-10
    k5 = est.cross_validate(folds=3)
    assert np.isfinite(k5.estimate).mean() > 0.9
