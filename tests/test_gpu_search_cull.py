"""The uncapped wave search skips a cell whose box lies beyond the K-th key
held (its distance taken from the target's place in its cell, in FP32 with
a margin). The neighbour sets must stay those of the reference's search
(search.rs:951-982, the oracle's tree walk) bit for bit: rows of a block
model and random targets, targets beyond the grid, repeated targets, targets
on samples (a K-th distance of zero), gridded samples with exact distance
ties, targets with fewer than K samples in reach, and a left-out sample."""
import numpy as np
import pytest

import oracle
from boitata_amd import _abi

pytestmark = pytest.mark.gpu


def check(engine, D, T, search, vg=None, hole=None, left=None):
    v = np.arange(len(D), dtype=np.float64)
    engine.set_samples(D[:, 0], D[:, 1], D[:, 2], v, None, hole)
    d = oracle.Data(D[:, 0], D[:, 1], D[:, 2], v, None, hole)
    gi, gc, gs = engine.neighbors(search, vg, T[:, 0], T[:, 1], T[:, 2], left_out=left)
    oi, oc, os_ = oracle.neighbors(d, search, vg, T, left_out=left)
    assert (gs == os_).all(), np.nonzero(gs != os_)[0][:10]
    assert (gc == oc).all(), np.nonzero(gc != oc)[0][:10]
    bad = [i for i in range(len(T)) if not (gi[i, :oc[i]] == oi[i, :oc[i]]).all()]
    assert not bad, bad[:10]
    return oc


def block_rows(n, lo, hi):
    axes = [np.linspace(lo[k], hi[k], n[k]) for k in range(3)]
    z, y, x = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
    return np.stack([x.ravel(), y.ravel(), z.ravel()], 1)  # x fastest


@pytest.mark.parametrize("K", [1, 8, 32, 48, 64])
def test_block_rows_and_random_targets(engine, K):
    rng = np.random.default_rng(K)
    D = rng.uniform(0, 100, (4000, 3))
    T = np.concatenate([block_rows((40, 10, 6), (0, 0, 0), (100, 100, 100)),
                        rng.uniform(-20, 120, (3000, 3))])
    search = _abi.make_search(1, K, 35.0, rotation=(30, 20, 10), ratios=(0.7, 0.4))
    oc = check(engine, D, T, search)
    assert (oc < K).any() and (oc == K).any()


def test_repeated_targets_on_samples_and_ties(engine):
    # Samples on a 1 m lattice: many exact distance ties; targets on samples
    # (K-th distance 0 for K = 1), on lattice midpoints, and repeated.
    g = np.arange(12.0)
    D = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    on = D[::7]
    mid = D[::11] + 0.5
    T = np.concatenate([on, np.repeat(mid, 3, axis=0), on[::-1]])
    for K in (1, 6, 26, 32, 40):
        check(engine, D, T, _abi.make_search(1, K, 3.0))
        check(engine, D, T, _abi.make_search(1, K, 50.0))


def test_sparse_reach_and_jumps(engine):
    rng = np.random.default_rng(3)
    D = np.concatenate([rng.uniform(0, 10, (300, 3)), rng.uniform(500, 510, (20, 3))])
    # Runs that alternate between the dense cluster, the sparse one (fewer
    # than K in reach) and empty space.
    T = np.concatenate([rng.uniform(0, 10, (200, 3)), rng.uniform(495, 515, (200, 3)),
                        rng.uniform(200, 300, (64, 3))])
    T = T[rng.permutation(len(T))[:300]]
    T = np.concatenate([T, np.sort(T, axis=0)])
    for K in (4, 32, 64):
        check(engine, D, T, _abi.make_search(1, K, 12.0))


def test_left_out_takes_no_bound(engine):
    rng = np.random.default_rng(4)
    D = rng.uniform(0, 50, (2000, 3))
    T = D[:1500].copy()
    left = np.arange(1500, dtype=np.int64)
    for K in (8, 32, 64):
        check(engine, D, T, _abi.make_search(1, K, 20.0), left=left)
