"""Capped searches whose candidate pool outgrows the 512 LDS entries between
two releases: the targets are searched again with a 65,536-entry pool in
device memory, so the lists equal the oracle's Selector
(crates/estimation/src/search.rs:476-577) and no target reports
BK_UNSUPPORTED. Coarse grid cells (cell factor 6) make every shell hold
thousands of candidates; the hole cap keeps the pool from being pruned."""
import numpy as np
import pytest

import oracle
from boitata_amd import _abi

pytestmark = pytest.mark.gpu


@pytest.fixture
def coarse():
    """A context of its own with cells six times the default side."""
    from boitata_amd.engine import Engine

    e = Engine(0)
    e.tuning(cell_factor=6.0)
    yield e
    e.close()


def dense_holes(seed, n=24000):
    rng = np.random.default_rng(seed)
    P = rng.normal(0, 15, (n, 3))
    v = rng.uniform(size=n)
    hole = (np.arange(n) // 40).astype(np.int32)
    return P, v, hole, rng


@pytest.mark.parametrize("kw", [
    dict(max_per_hole=2),
    dict(max_per_hole=1, octant=True, max_per_octant=3),
    dict(sectors=6, max_per_sector=4, max_per_hole=3),
])
def test_overflowing_pools_give_the_oracle_lists(coarse, kw):
    P, v, hole, rng = dense_holes(31)
    T = rng.normal(0, 15, (300, 3))
    s = _abi.make_search(1, 32, 60.0, **kw)
    coarse.set_samples(P[:, 0], P[:, 1], P[:, 2], v, None, hole)
    d = oracle.Data(P[:, 0], P[:, 1], P[:, 2], v, None, hole)
    gi, gc, gs = coarse.neighbors(s, None, T[:, 0], T[:, 1], T[:, 2])
    retried = coarse.stat(_abi.STAT_SEARCH_RETRY)
    oi, oc, os_ = oracle.neighbors(d, s, None, T)
    print(f"{kw}: {retried} of {len(T)} targets searched again with the device-memory pool")
    assert retried > 0
    assert not (gs == _abi.UNSUPPORTED).any()
    assert (gs == os_).all() and (gc == oc).all()
    for i in range(len(T)):
        assert (gi[i, :oc[i]] == oi[i, :oc[i]]).all(), i


def test_overflowing_pools_in_a_kriging_pass(coarse):
    P, v, hole, rng = dense_holes(32)
    T = rng.normal(0, 15, (200, 3))
    s = _abi.make_search(4, 24, 60.0, max_per_hole=2)
    vg = _abi.make_variogram([("spherical", 0.9, 40.0)], nugget=0.1)
    coarse.set_samples(P[:, 0], P[:, 1], P[:, 2], v, None, hole)
    d = oracle.Data(P[:, 0], P[:, 1], P[:, 2], v, None, hole)
    g = coarse.krige(vg, s, T[:, 0], T[:, 1], T[:, 2], precision=_abi.FP64)
    retried = coarse.stat(_abi.STAT_SEARCH_RETRY)
    o = oracle.krige(d, vg, s, T)
    assert retried > 0
    assert (g["status"] == o["status"]).all()
    ok = o["status"] == 0
    assert ok.all()
    err = np.abs(g["value"] - o["value"]) / np.maximum(np.abs(o["value"]), 1e-3)
    assert err.max() <= 1e-10
