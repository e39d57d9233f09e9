"""bk_sgs (sequential Gaussian simulation in scores) against the oracle's
sequential loop (or_sgs, crates/simulation/src/sgs.rs node loop), node for
node, through the C ABI: same path and draws, the same source per node (datum
index, kriged, marginal) and scores within 1e-9 (FP64) or 1e-6 (FP32 + FP64
refinement) relative with a floor of 1e-3 (the scores have unit variance).
Cases: drill-hole data under a block grid with a rotated nested model and
ellipsoid search; octants (the sector kernel with index limits); a hole cap
(the pool search); planar data under a grid that leaves the plane; nodes on
data and out of reach; the plan form at scale (50k nodes) against the
oracle's plan form."""
import numpy as np
import pytest

import oracle
from boitata_amd import _abi, datasets
from boitata_amd.engine import EngineError

pytestmark = pytest.mark.gpu


def scores_of(v):
    """Normal scores of the values by rank (ties by index): test data only."""
    from scipy.stats import norm

    r = np.argsort(np.argsort(v, kind="stable"), kind="stable")
    return norm.ppf((r + 0.5) / len(v))


def run_both(engine, D, s, G, vg, search, seed, hole=None, plan=False, precisions=(_abi.FP64, _abi.FP32_IR)):
    rng = np.random.default_rng(seed)
    path = rng.permutation(len(G))
    z = rng.normal(size=len(G))
    engine.set_samples(D[:, 0], D[:, 1], D[:, 2], s, None, hole)
    d = oracle.Data(D[:, 0], D[:, 1], D[:, 2], s, None, hole)
    want, wsrc = oracle.sgs(d, vg, search, G, path, z, plan=plan)
    out = {}
    for prec in precisions:
        g = engine.sgs(vg, search, G[:, 0], G[:, 1], G[:, 2], path, z, precision=prec)
        assert (g["source"] == wsrc).all(), np.nonzero(g["source"] != wsrc)[0][:10]
        tol = 1e-9 if prec == _abi.FP64 else 1e-6
        err = np.abs(g["score"] - want) / np.maximum(np.abs(want), 1e-3)
        assert err.max() <= tol, (prec, err.max())
        out[prec] = (g, err.max())
    return want, wsrc, out


def block_grid(n, lo, hi):
    axes = [np.linspace(lo[k], hi[k], n[k]) for k in range(3)]
    return np.stack(np.meshgrid(*axes, indexing="ij"), -1).reshape(-1, 3)


def test_drillholes_under_a_block_grid(engine):
    w = datasets.config1()
    D = np.stack([w.x, w.y, w.z], 1)[::3]
    s = scores_of(w.value[::3])
    G = block_grid((24, 24, 8), (20, 20, -190), (480, 480, -10))
    G[:50] = D[:50]  # nodes on data
    vg = _abi.make_variogram([("exponential", 0.6, 150.0), ("spherical", 0.3, 300.0)], nugget=0.1,
                             rotation=(30, 0, 0), ratios=(0.6, 0.3))
    search = _abi.make_search(4, 20, 200.0, rotation=(30, 0, 0), ratios=(0.6, 0.3))
    want, src, out = run_both(engine, D, s, G, vg, search, 1)
    assert (src[:50] == np.arange(50)).all()
    print(f"drill holes: {len(G)} nodes, max rel err FP64 {out[_abi.FP64][1]:.2e}, "
          f"FP32+IR {out[_abi.FP32_IR][1]:.2e}, draw passes {engine.stat(_abi.STAT_SGS_PASSES)}")


@pytest.mark.parametrize("kw", [dict(octant=True), dict(octant=True, max_per_octant=3), dict(max_per_hole=2)])
def test_capped_searches(engine, kw):
    rng = np.random.default_rng(2)
    D = rng.uniform(0, 100, (1500, 3))
    s = rng.normal(size=len(D))
    hole = (np.arange(len(D)) // 8).astype(np.int32)
    G = block_grid((20, 20, 10), (2, 2, 2), (98, 98, 98))
    vg = _abi.make_variogram([("spherical", 1.0, 40.0)], nugget=0.2)
    search = _abi.make_search(2, 24, 45.0, **kw)
    run_both(engine, D, s, G, vg, search, 3, hole=hole)


def test_planar_data_and_nodes_off_the_plane(engine):
    rng = np.random.default_rng(4)
    D = np.concatenate([rng.uniform(0, 100, (600, 2)), np.zeros((600, 1))], 1)
    s = rng.normal(size=len(D))
    G = block_grid((30, 30, 3), (1, 1, 0), (99, 99, 10))
    vg = _abi.make_variogram([("gaussian", 1.0, 35.0)], nugget=0.05)
    for kw in (dict(octant=True), dict()):
        run_both(engine, D, s, G, vg, _abi.make_search(1, 16, 40.0, **kw), 5)


def test_out_of_reach_and_sparse_nodes(engine):
    rng = np.random.default_rng(6)
    D = rng.uniform(0, 50, (200, 3))
    s = rng.normal(size=len(D))
    G = np.concatenate([rng.uniform(0, 50, (500, 3)), rng.uniform(500, 600, (100, 3))])
    vg = _abi.make_variogram([("exponential", 1.0, 20.0)])
    want, src, _ = run_both(engine, D, s, G, vg, _abi.make_search(1, 12, 15.0), 7)
    assert (src == -2).sum() > 0


def test_plan_form_at_scale(engine):
    w = datasets.config2(holes=(10, 10), targets=False)
    D = np.stack([w.x, w.y, w.z], 1)
    s = scores_of(w.value)
    G = block_grid((50, 50, 20), (10, 10, -490), (990, 990, -10))
    vg = _abi.make_variogram([("exponential", 0.5, 150.0), ("spherical", 0.4, 400.0)], nugget=0.1,
                             rotation=(30, 20, 0), ratios=(250 / 400, 80 / 400))
    search = _abi.make_search(4, 32, 400.0, rotation=(30, 20, 0), ratios=(250 / 400, 80 / 400))
    _, _, out = run_both(engine, D, s, G, vg, search, 8, plan=True)
    print(f"{len(G)} nodes: FP64 {out[_abi.FP64][1]:.2e}, FP32+IR {out[_abi.FP32_IR][1]:.2e}, "
          f"passes {engine.stat(_abi.STAT_SGS_PASSES)}")


def test_invalid_inputs(engine):
    rng = np.random.default_rng(9)
    D = rng.uniform(0, 10, (30, 3))
    engine.set_samples(D[:, 0], D[:, 1], D[:, 2], rng.normal(size=30))
    G = rng.uniform(0, 10, (20, 3))
    vg = _abi.make_variogram([("spherical", 1.0, 5.0)])
    s = _abi.make_search(1, 8, 10.0)
    with pytest.raises(EngineError):
        engine.sgs(vg, s, G[:, 0], G[:, 1], G[:, 2], np.zeros(20, np.int64), np.zeros(20))
    G[3] = G[4]
    with pytest.raises(EngineError):
        engine.sgs(vg, s, G[:, 0], G[:, 1], G[:, 2], np.arange(20), np.zeros(20))


def test_host_class_realizations():
    from boitata_amd.estimation import Search, Variogram
    from boitata_amd.simulation import SGS

    rng = np.random.default_rng(10)
    D = rng.uniform(0, 200, (800, 3))
    s = rng.normal(size=len(D))
    G = block_grid((16, 16, 6), (5, 5, 5), (195, 195, 195))
    vg = Variogram([("spherical", 0.8, 90.0)], nugget=0.2)
    search = Search(ranges=(90.0, 90.0, 90.0), min_samples=2, max_samples=16)
    sim = SGS(vg, search).fit(D, s)
    got = sim.simulate(G, n=2, seed=3)
    d = oracle.Data(D[:, 0], D[:, 1], D[:, 2], s)
    for r in range(2):
        g = np.random.default_rng(3 + r)
        path, z = g.permutation(len(G)), g.normal(size=len(G))
        want, _ = oracle.sgs(d, vg._c, search._c, G, path, z)
        assert np.max(np.abs(got[r] - want) / np.maximum(np.abs(want), 1e-3)) <= 1e-9
    assert not np.array_equal(got[0], got[1])
    sim.close()
