"""Parity of the HIP path with the oracle, through the C ABI.

Neighbour sets are compared bit for bit (indices in selection order);
estimates and variances at 1e-10 relative for the FP64 form and 1e-6 for
FP32 + refinement, relative with a floor of 1e-3 * sill (the floor of the
reference's own close(), crates/estimation/src/kernel.rs:178-180)."""
import numpy as np
import pytest

import oracle
from boitata_amd import _abi, datasets

pytestmark = pytest.mark.gpu


def sill_of(vg):
    return vg.nugget + sum(vg.structures[k].sill for k in range(vg.n_structures))


def assert_neighbours(engine, d, search, vg, T, left=None):
    gi, gc, gs = engine.neighbors(search, vg, T[:, 0], T[:, 1], T[:, 2], left_out=left)
    oi, oc, os_ = oracle.neighbors(d, search, vg, T, left_out=left)
    assert (gs == os_).all(), np.nonzero(gs != os_)[0][:10]
    assert (gc == oc).all()
    for i in range(len(T)):
        assert (gi[i, :oc[i]] == oi[i, :oc[i]]).all(), i


def assert_estimates(g, o, vg, tol, keys=("value", "variance")):
    assert (g["status"] == o["status"]).all(), (np.bincount(g["status"]), np.bincount(o["status"]))
    ok = o["status"] == 0
    floor = 1e-3 * sill_of(vg)
    for k in keys:
        a, b = np.asarray(g[k])[ok], np.asarray(o[k])[ok]
        err = np.abs(a - b) / np.maximum(np.abs(b), floor)
        assert err.max(initial=0.0) <= tol, (k, err.max())
    assert np.isnan(np.asarray(g["value"])[~ok]).all()


@pytest.fixture(scope="module")
def c1_data():
    w = datasets.config1()
    return w, oracle.Data(w.x, w.y, w.z, w.value, None, w.hole), w.targets[::23].copy()


@pytest.fixture
def c1(engine, c1_data):
    """Config 1, its samples loaded on the engine (other tests load others)."""
    w = c1_data[0]
    engine.set_samples(w.x, w.y, w.z, w.value, None, w.hole)
    return c1_data


def test_c1_neighbour_sets_bit_identical(engine, c1):
    w, d, T = c1
    engine.set_samples(w.x, w.y, w.z, w.value, None, w.hole)
    assert_neighbours(engine, d, w.search, w.variogram, T)


@pytest.mark.parametrize("kw", [
    dict(octant=True),
    dict(octant=True, max_per_octant=2, min_sectors=4, max_empty_sectors=2),
    dict(max_per_hole=3),
    dict(max_per_hole=2, octant=True, min_holes=3),
    dict(sectors=6, max_per_sector=3, min_sectors=3),
    dict(min_holes=3),
])
def test_capped_searches_bit_identical(engine, c1, kw):
    w, d, T = c1
    s = _abi.make_search(1, 24, 120.0, rotation=(30, 0, 0), ratios=(80 / 120, 30 / 120), **kw)
    assert_neighbours(engine, d, s, w.variogram, T[:1200])


@pytest.mark.parametrize("mode", [-1, 1, 2, 3])
def test_search_modes_agree(engine, c1, mode):
    # Lane per target, and tiles of several reaches (reach 1 hands many
    # targets back to the lane walk): every mode gives the oracle's sets.
    w, d, T = c1
    try:
        engine.search_tuning(mode)
        assert_neighbours(engine, d, w.search, w.variogram, T)
        s = _abi.make_search(1, 64, 200.0, rotation=(30, 0, 0), ratios=(80 / 120, 30 / 120))
        assert_neighbours(engine, d, s, w.variogram, T[:3000])
        if mode == 1:
            assert engine.stat() > 0
    finally:
        engine.search_tuning(0)


def test_wave_search_hands_sparse_targets_back(engine):
    # Targets far outside the data: the shell table ends before K are found,
    # the lane walk continues on rings.
    rng = np.random.default_rng(5)
    P = rng.uniform(0, 100, (2000, 3))
    v = rng.uniform(size=2000)
    engine.set_samples(P[:, 0], P[:, 1], P[:, 2], v)
    d = oracle.Data(P[:, 0], P[:, 1], P[:, 2], v)
    T = np.concatenate([rng.uniform(0, 100, (500, 3)), rng.uniform(2000, 3000, (64, 3))])
    s = _abi.make_search(1, 16, 1e5)
    assert_neighbours(engine, d, s, None, T)
    assert engine.stat() > 0


def test_leave_one_out_neighbours(engine, c1):
    w, d, _ = c1
    idx = np.arange(0, w.n, 7)
    P = np.stack([w.x, w.y, w.z], 1)[idx]
    assert_neighbours(engine, d, w.search, w.variogram, P, left=idx)


@pytest.mark.parametrize("kind,param", [(_abi.ORDINARY, 0.0), (_abi.SIMPLE, 1.0), (_abi.INDICATOR, 1.2)])
@pytest.mark.parametrize("precision,tol", [(_abi.FP64, 1e-10), (_abi.FP32_IR, 1e-6)])
def test_c1_kriging(engine, c1, kind, param, precision, tol):
    w, d, T = c1
    g = engine.krige(w.variogram, w.search, T[:, 0], T[:, 1], T[:, 2], kind, param, precision)
    o = oracle.krige(d, w.variogram, w.search, T, kind, param)
    assert_estimates(g, o, w.variogram, tol, ("value", "variance", "lagrange"))
    assert (g["n_used"][o["status"] == 0] == o["n_used"][o["status"] == 0]).all()


def test_c2_per_structure_ranges(engine):
    w = datasets.config2(holes=(8, 8), target_counts=(40, 40, 10))
    engine.set_samples(w.x, w.y, w.z, w.value, None, w.hole)
    d = oracle.Data(w.x, w.y, w.z, w.value, None, w.hole)
    T = w.targets[::3].copy()
    assert_neighbours(engine, d, w.search, w.variogram, T[:2000])
    for vg in (w.variogram, w.variogram_shared):
        o = oracle.krige(d, vg, w.search, T)
        for prec, tol in ((_abi.FP64, 1e-10), (_abi.FP32_IR, 1e-6)):
            assert_estimates(engine.krige(vg, w.search, T[:, 0], T[:, 1], T[:, 2], precision=prec), o, vg, tol)


def test_c3_clustered_octants(engine):
    w = datasets.config3(target_counts=(40, 40, 20))
    engine.set_samples(w.x, w.y, w.z, w.value)
    d = oracle.Data(w.x, w.y, w.z, w.value)
    T = w.targets[::5].copy()
    assert_neighbours(engine, d, w.search, w.variogram, T)
    o = oracle.krige(d, w.variogram, w.search, T)
    for prec, tol in ((_abi.FP64, 1e-10), (_abi.FP32_IR, 1e-6)):
        g = engine.krige(w.variogram, w.search, T[:, 0], T[:, 1], T[:, 2], precision=prec)
        assert_estimates(g, o, w.variogram, tol)
        counts = np.bincount(g["n_used"], minlength=65)
        assert counts[:16].sum() > 0 and counts[33:].sum() > 0  # uneven neighbourhoods
    # Every refine limit: the buckets above it take the FP64 kernel.
    try:
        for limit in (16, 64, 32):
            engine.refine_limit(limit)
            g = engine.krige(w.variogram, w.search, T[:, 0], T[:, 1], T[:, 2], precision=_abi.FP32_IR)
            assert_estimates(g, o, w.variogram, 1e-6)
    finally:
        engine.refine_limit(32)


def test_edge_cases(engine):
    rng = np.random.default_rng(8)
    P = rng.uniform(0, 100, (300, 3))
    v = rng.uniform(0, 5, 300)
    err = np.where(rng.uniform(size=300) < 0.2, 0.3, 0.0)
    # Coincident samples (merged, coincident.rs) and near ones.
    P[10] = P[11]
    P[20] = P[21] + 1e-8
    P[30] = P[31] + 1e-3
    vg = _abi.make_variogram([("spherical", 0.8, 60.0)], nugget=0.2, rotation=(20, 10, 5), ratios=(0.7, 0.5))
    engine.set_samples(P[:, 0], P[:, 1], P[:, 2], v, err, None)
    d = oracle.Data(P[:, 0], P[:, 1], P[:, 2], v, err)
    T = np.concatenate([P[[10, 20, 30]] + 0.5, rng.uniform(-20, 120, (300, 3))])
    for s in (_abi.make_search(1, 12, 40.0), _abi.make_search(6, 64, 80.0), _abi.make_search(1, 70, 200.0),
              _abi.make_search(10, 16, 15.0)):
        o = oracle.krige(d, vg, s, T)
        for prec, tol in ((_abi.FP64, 1e-10), (_abi.FP32_IR, 1e-6)):
            assert_estimates(engine.krige(vg, s, T[:, 0], T[:, 1], T[:, 2], precision=prec), o, vg, tol)
    # A model without a sill goes to the bordered LU.
    pw = _abi.make_variogram([("power", 1.0, 50.0, 1.5)], nugget=0.1)
    s = _abi.make_search(1, 16, 60.0)
    o = oracle.krige(d, pw, s, T)
    assert_estimates(engine.krige(pw, s, T[:, 0], T[:, 1], T[:, 2]), o, pw, 1e-9)


def test_no_samples_and_invalid_parameters(engine):
    engine.set_samples(np.zeros(0), np.zeros(0), np.zeros(0), np.zeros(0))
    vg = _abi.make_variogram([("spherical", 1.0, 10.0)])
    g = engine.krige(vg, _abi.make_search(1, 4, 10.0), np.zeros(3), np.zeros(3), np.zeros(3))
    assert (g["status"] == _abi.SEARCH_FAILED).all() and np.isnan(g["value"]).all()
    from boitata_amd.engine import EngineError

    with pytest.raises(EngineError):
        engine.krige(vg, _abi.make_search(5, 4, 10.0), np.zeros(1), np.zeros(1), np.zeros(1))
    with pytest.raises(EngineError):
        engine.krige(_abi.make_variogram([("power", 1.0, 10.0, 1.0)]), _abi.make_search(1, 4, 10.0), np.zeros(1),
                     np.zeros(1), np.zeros(1), kind=_abi.SIMPLE)


def lanes_of(count, edges=(16, 32, 64)):
    """The lane group (size bucket) of each system under the default edges."""
    count = np.asarray(count)
    return np.where(count <= edges[0], edges[0], np.where(count <= edges[1], edges[1], edges[2]))


def test_block_kriging(engine):
    w = datasets.config5(target_counts=(40, 40, 10), holes=(16, 16))
    engine.tuning(edges=(16, 32, 64))
    engine.set_samples(w.x, w.y, w.z, w.value, None, w.hole)
    d = oracle.Data(w.x, w.y, w.z, w.value, None, w.hole)
    T = w.targets[::7].copy()
    o = oracle.block_krige(d, w.variogram, w.search, w.block, T)
    # Gaussian covariance with a 1e-4 nugget and 1 cm twins. The device's FP64
    # blocks follow the oracle's engine-order path (Cholesky, Schur scalar) to
    # 1e-10; against the reference's LU the gap is bounded by kappa * eps.
    idx, cnt, st = oracle.neighbors(d, w.search, w.variogram, T)
    eng = oracle.block_krige_lists(d, w.variogram, w.block, T, idx, cnt, solver=1, lanes=lanes_of(cnt))
    eng["status"] = np.where(st != 0, st, eng["status"])
    g = engine.block_krige(w.variogram, w.search, w.block, T[:, 0], T[:, 1], T[:, 2], precision=_abi.FP64)
    assert_estimates(g, eng, w.variogram, 1e-10, ("value", "variance", "support"))
    ok = o["status"] == 0
    kappa = np.array([np.linalg.cond(oracle.system_matrix(d, w.variogram, idx[i, :cnt[i]])) if ok[i] else 1.0
                      for i in range(len(T))])
    floor = 1e-3 * sill_of(w.variogram)
    for k in ("value", "variance"):
        gap = np.abs(g[k][ok] - o[k][ok]) / np.maximum(np.abs(o[k][ok]), floor)
        assert (gap <= 1e-10 + 64 * kappa[ok] * np.finfo(float).eps).all(), (k, gap.max(), kappa.max())
    g32 = engine.block_krige(w.variogram, w.search, w.block, T[:, 0], T[:, 1], T[:, 2], precision=_abi.FP32_IR)
    assert_estimates(g32, o, w.variogram, 1e-6, ("value", "variance", "support"))


@pytest.mark.parametrize("max_samples", [80, 128, 256])
def test_systems_beyond_the_cholesky_kernel(engine, c1, max_samples):
    # Searches of up to 256 samples: systems above 64 take the bordered LU
    # (krige.rs:127-201), above 79 with its matrix in device memory. Point and
    # block support, both precisions, ordinary and simple kriging.
    w, d, T = c1
    engine.set_samples(w.x, w.y, w.z, w.value, None, w.hole)
    s = _abi.make_search(4, max_samples, 400.0, rotation=(30, 0, 0), ratios=(80 / 120, 30 / 120))
    S = T[::6]
    for kind, param in ((_abi.ORDINARY, 0.0), (_abi.SIMPLE, 1.0)):
        o = oracle.krige(d, w.variogram, s, S, kind, param)
        assert (o["n_used"][o["status"] == 0] > 64).mean() > 0.5
        for prec in (_abi.FP64, _abi.FP32_IR):
            g = engine.krige(w.variogram, s, S[:, 0], S[:, 1], S[:, 2], kind, param, prec)
            assert not (g["status"] == _abi.UNSUPPORTED).any()
            assert_estimates(g, o, w.variogram, 1e-10, ("value", "variance", "lagrange"))
    blk = _abi.make_block((10, 10, 10), (2, 2, 2))
    o = oracle.block_krige(d, w.variogram, s, blk, S[::4])
    g = engine.block_krige(w.variogram, s, blk, S[::4, 0], S[::4, 1], S[::4, 2])
    assert_estimates(g, o, w.variogram, 1e-10, ("value", "variance", "support"))


def test_block_kriging_well_conditioned(engine, c1):
    w, d, T = c1
    engine.set_samples(w.x, w.y, w.z, w.value, None, w.hole)
    blk = _abi.make_block((10, 10, 10), (4, 4, 2), rotation=(15, 0, 0))
    o = oracle.block_krige(d, w.variogram, w.search, blk, T)
    for prec, tol in ((_abi.FP64, 1e-10), (_abi.FP32_IR, 1e-6)):
        g = engine.block_krige(w.variogram, w.search, blk, T[:, 0], T[:, 1], T[:, 2], precision=prec)
        assert_estimates(g, o, w.variogram, tol, ("value", "variance", "support"))


def test_leave_one_out(engine, c1):
    w, d, _ = c1
    engine.set_samples(w.x, w.y, w.z, w.value, None, w.hole)
    which = np.arange(0, w.n, 3)
    o = oracle.loo(d, w.variogram, w.search, which)
    for prec, tol in ((_abi.FP64, 1e-10), (_abi.FP32_IR, 1e-6)):
        g = engine.loo(w.variogram, w.search, which, precision=prec)
        assert_estimates(g, o, w.variogram, tol)


@pytest.mark.parametrize("edges", [(16, 32, 64), (32, 32, 64), (64, 64, 64)])
def test_refined_form_keeps_well_conditioned_systems(engine, c1, edges):
    """Every system of the CPU-runnable workload (nugget 0.2 of a unit sill) is
    solved by refinement, in one sweep, by each refined kernel (bucket edges
    16 / 32 / 64): none is handed back to the FP64 kernel, which would give
    correct estimates at the FP64 kernel's cost."""
    w, d, T = c1
    engine.tuning(edges=edges)
    engine.refine_limit(64)
    try:
        g = engine.krige(w.variogram, w.search, T[:, 0], T[:, 1], T[:, 2], precision=_abi.FP32_IR)
        solved = int((g["status"] == 0).sum())
        assert engine.stat(_abi.STAT_REFINE_REDO) == 0
        assert engine.stat(_abi.STAT_REFINED) == solved
        assert engine.stat(_abi.STAT_REFINE_SWEEPS) <= 1.01 * solved
    finally:
        engine.tuning(edges=(16, 32, 64))


def test_results_are_deterministic(engine, c1):
    w, d, T = c1
    engine.set_samples(w.x, w.y, w.z, w.value, None, w.hole)
    a = engine.krige(w.variogram, w.search, T[:, 0], T[:, 1], T[:, 2], precision=_abi.FP32_IR)
    engine.tuning(cell_factor=1.7)
    b = engine.krige(w.variogram, w.search, T[:, 0], T[:, 1], T[:, 2], precision=_abi.FP32_IR)
    engine.tuning(cell_factor=0.65)
    assert (a["value"].view(np.uint64) == b["value"].view(np.uint64)).all()
    assert (a["variance"].view(np.uint64) == b["variance"].view(np.uint64)).all()


@pytest.mark.parametrize("prec", [_abi.FP64, _abi.FP32_IR])
def test_autotune_keeps_results(engine, c1, prec):
    """Tuning never changes neighbour sets or statuses. FP64 results are the
    same bit for bit; FP32+IR results agree to the refinement's tolerance
    (the bucket edges and refine limit pick the refined kernel of a size
    class, whose FP32 eliminations differ), and bit for bit when the tuned
    edges and refine limit are the defaults."""
    w, d, T = c1
    engine.refine_limit(64)
    before = engine.krige(w.variogram, w.search, T[:, 0], T[:, 1], T[:, 2], precision=prec)
    rep = engine.autotune(w.variogram, w.search, T[:, 0], T[:, 1], T[:, 2], precision=prec, sample=1500)
    assert rep["waves"] in (1, 2, 4) and rep["search_mode"] in (-1, 0, 2)
    assert rep["search_ms_after"] <= rep["search_ms_before"] and rep["solve_ms_after"] <= rep["solve_ms_before"]
    after = engine.krige(w.variogram, w.search, T[:, 0], T[:, 1], T[:, 2], precision=prec)
    same_kernels = (rep["edge0"], rep["edge1"], rep["edge2"], rep["refine_limit"]) == (16, 32, 64, 64)
    exact = ("value", "variance", "status", "n_used") if prec == _abi.FP64 or same_kernels else ("status", "n_used")
    for k in exact:
        assert (np.asarray(before[k]).view(np.uint8) == np.asarray(after[k]).view(np.uint8)).all(), k
    assert_estimates(after, before, w.variogram, 1e-6)
    # Back to the defaults for the tests that follow.
    engine.tuning(edges=(16, 32, 64), cell_factor=0.65)
    engine.search_tuning(0)
    engine.solve_tuning(2)
    engine.refine_limit(64)


def test_graph_replay_matches_eager(engine, c1):
    """Repeated passes over device arrays run as a captured graph: every replay
    equals the ordinary pass bit for bit, with other work and synchronisations
    in between, and new samples or tuning take effect (a new capture)."""
    import torch

    w, d, T = c1
    dev = torch.device("cuda:0")
    tx, ty, tz = (torch.from_numpy(np.ascontiguousarray(T[:, k])).to(dev) for k in range(3))
    keys = ("value", "variance", "lagrange", "n_used", "status")

    def bits(out):
        return {k: out[k].cpu().numpy().view(np.uint8).copy() for k in keys}

    engine.graphs(False)
    ref = bits(engine.krige(w.variogram, w.search, tx, ty, tz, precision=_abi.FP32_IR))
    engine.graphs(True)
    out = engine.krige(w.variogram, w.search, tx, ty, tz, precision=_abi.FP32_IR)
    start = engine.stat(_abi.STAT_GRAPH_LAUNCHES)
    for rep in range(5):
        for k in keys:
            out[k].fill_(-1)
        engine.krige(w.variogram, w.search, tx, ty, tz, precision=_abi.FP32_IR, out=out)
        got = bits(out)
        for k in keys:
            assert (got[k] == ref[k]).all(), (rep, k)
        # Other work on the context and on torch's stream between replays.
        engine.krige(w.variogram, w.search, T[:50, 0], T[:50, 1], T[:50, 2], precision=_abi.FP64)
        torch.cuda.synchronize()
    assert engine.stat(_abi.STAT_GRAPH_LAUNCHES) - start >= 3
    # New sample values: the captured pass must not be reused.
    engine.set_samples(w.x, w.y, w.z, 2.0 * w.value, None, w.hole)
    for _ in range(3):
        engine.krige(w.variogram, w.search, tx, ty, tz, precision=_abi.FP32_IR, out=out)
    scaled = out["value"].cpu().numpy()
    ok = ref["status"].view(np.int32) == 0
    assert np.allclose(scaled[ok], 2.0 * ref["value"].view(np.float64)[ok], rtol=1e-12, atol=0.0)
    engine.set_samples(w.x, w.y, w.z, w.value, None, w.hole)
    # A tuning change (workgroup size) is a new capture with the same results.
    engine.solve_tuning(1)
    for _ in range(3):
        engine.krige(w.variogram, w.search, tx, ty, tz, precision=_abi.FP32_IR, out=out)
        got = bits(out)
        for k in keys:
            assert (got[k] == ref[k]).all(), k
    engine.solve_tuning(2)


@pytest.mark.parametrize("model,param", [("cubic", 0.0), ("pentaspherical", 0.0), ("circular", 0.0),
                                         ("sinehole", 0.0), ("matern", 0.5), ("matern", 1.5), ("matern", 2.5)])
def test_every_stationary_model(engine, c1, model, param):
    # model.rs:76-126: each stationary shape, nested with a spherical
    # structure under the workload's anisotropy, point and block support, in
    # both precisions (the refined form takes cubic and pentaspherical; the
    # others run the FP64 kernel in either precision).
    w, d, T = c1
    engine.set_samples(w.x, w.y, w.z, w.value, None, w.hole)
    vg = _abi.make_variogram([(model, 0.5, 90.0, param), ("spherical", 0.3, 200.0)], nugget=0.2,
                             rotation=(30, 0, 0), ratios=(80 / 120, 30 / 120))
    S = T[::3]
    o = oracle.krige(d, vg, w.search, S)
    assert (o["status"] == 0).mean() > 0.9
    for prec, tol in ((_abi.FP64, 1e-10), (_abi.FP32_IR, 1e-6)):
        g = engine.krige(vg, w.search, S[:, 0], S[:, 1], S[:, 2], precision=prec)
        assert_estimates(g, o, vg, tol, ("value", "variance", "lagrange"))
    blk = _abi.make_block((10, 10, 10), (2, 2, 2))
    o = oracle.block_krige(d, vg, w.search, blk, S[::5])
    for prec, tol in ((_abi.FP64, 1e-10), (_abi.FP32_IR, 1e-6)):
        g = engine.block_krige(vg, w.search, blk, S[::5, 0], S[::5, 1], S[::5, 2], precision=prec)
        assert_estimates(g, o, vg, tol, ("value", "variance", "support"))
