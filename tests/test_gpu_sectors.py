"""Octant and quadrant searches without a hole cap (the sector-list kernel,
with k_search for the targets it hands back) against the oracle's Selector
(crates/estimation/src/search.rs:476-577), neighbour lists bit for bit:
- samples on a lattice shared with the targets, unrotated, so that many
  offsets lie exactly on a sector plane (the sign test defers to the world
  computation there);
- planar data (quadrants, up to 16 per quadrant);
- max_samples below sectors x cap (the selection stops at full);
- caps beyond the kernel's lists (k_search alone);
- leave-one-out, duplicated samples, far targets handed back, rotated
  ellipsoids with dip and rake, and the sector rules."""
import numpy as np
import pytest

import oracle
from boitata_amd import _abi

pytestmark = pytest.mark.gpu


def check(engine, P, v, search, T, left=None, hole=None):
    engine.set_samples(P[:, 0], P[:, 1], P[:, 2], v, None, hole)
    d = oracle.Data(P[:, 0], P[:, 1], P[:, 2], v, None, hole)
    gi, gc, gs = engine.neighbors(search, None, T[:, 0], T[:, 1], T[:, 2], left_out=left)
    oi, oc, os_ = oracle.neighbors(d, search, None, T, left_out=left)
    assert (gs == os_).all(), np.nonzero(gs != os_)[0][:10]
    assert (gc == oc).all(), np.nonzero(gc != oc)[0][:10]
    for i in range(len(T)):
        assert (gi[i, :oc[i]] == oi[i, :oc[i]]).all(), (i, gi[i, :oc[i]], oi[i, :oc[i]])
    return oc


def lattice(n, step=1.0):
    g = np.arange(n) * step
    return np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)


def test_lattice_offsets_on_sector_planes(engine):
    P = lattice(14)
    v = np.arange(len(P), dtype=np.float64)
    T = P[::37] + np.array([0.0, 0.0, 0.0])
    T = np.concatenate([T, P[::53] + np.array([0.5, 0.0, 0.0])])
    for kw in (dict(octant=True), dict(octant=True, max_per_octant=8), dict(octant=True, max_per_octant=5)):
        s = _abi.make_search(1, 64, 6.0, **kw)
        check(engine, P, v, s, T)


def test_rotated_ellipsoid_with_dip_and_rake(engine):
    rng = np.random.default_rng(21)
    P = rng.uniform(0, 200, (6000, 3))
    v = rng.uniform(size=len(P))
    T = rng.uniform(0, 200, (1500, 3))
    for kw in (dict(octant=True), dict(octant=True, max_per_octant=4, min_sectors=6, max_empty_sectors=1)):
        s = _abi.make_search(2, 48, 60.0, rotation=(30, 20, 10), ratios=(0.6, 0.25), **kw)
        check(engine, P, v, s, T)


def test_planar_quadrants(engine):
    rng = np.random.default_rng(22)
    P = np.concatenate([rng.uniform(0, 100, (3000, 2)), np.full((3000, 1), 5.0)], 1)
    v = rng.uniform(size=len(P))
    T = np.concatenate([rng.uniform(0, 100, (800, 2)), np.full((800, 1), 5.0)], 1)
    for cap in (3, 12, 16):
        s = _abi.make_search(1, 64, 40.0, octant=True, max_per_octant=cap, rotation=(40, 0, 0), ratios=(0.5, 1.0))
        check(engine, P, v, s, T)


def test_selection_full_before_every_sector(engine):
    rng = np.random.default_rng(23)
    P = rng.uniform(0, 100, (5000, 3))
    v = rng.uniform(size=len(P))
    T = rng.uniform(0, 100, (1000, 3))
    for K, cap in ((20, 6), (7, 8), (33, 5)):
        s = _abi.make_search(1, K, 50.0, octant=True, max_per_octant=cap)
        oc = check(engine, P, v, s, T)
        assert oc.max() == K


def test_caps_beyond_the_lists(engine):
    rng = np.random.default_rng(24)
    P = rng.uniform(0, 100, (5000, 3))
    v = rng.uniform(size=len(P))
    T = rng.uniform(0, 100, (400, 3))
    s = _abi.make_search(1, 96, 60.0, octant=True, max_per_octant=12)
    check(engine, P, v, s, T)


def test_leave_one_out_and_duplicates(engine):
    rng = np.random.default_rng(25)
    P = rng.uniform(0, 50, (3000, 3))
    P[100:200] = P[0:100]  # exact twins: ties broken by index
    v = rng.uniform(size=len(P))
    which = np.arange(0, len(P), 5)
    s = _abi.make_search(1, 32, 20.0, octant=True)
    check(engine, P, v, s, P[which], left=which)
    check(engine, P, v, s, P[:300] + 1e-3)


def test_far_targets_handed_back(engine):
    rng = np.random.default_rng(26)
    P = rng.uniform(0, 100, (2000, 3))
    v = rng.uniform(size=len(P))
    T = np.concatenate([rng.uniform(0, 100, (300, 3)), rng.uniform(1500, 2500, (64, 3))])
    s = _abi.make_search(1, 16, 1e5, octant=True, max_per_octant=2)
    check(engine, P, v, s, T)
    assert engine.stat() > 0


def test_min_holes_rule_with_octants(engine):
    rng = np.random.default_rng(27)
    P = rng.uniform(0, 100, (3000, 3))
    v = rng.uniform(size=len(P))
    hole = (np.arange(len(P)) // 10).astype(np.int32)
    T = rng.uniform(0, 100, (500, 3))
    s = _abi.make_search(1, 24, 40.0, octant=True, min_holes=4)
    check(engine, P, v, s, T, hole=hole)
