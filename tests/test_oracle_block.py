"""The oracle's block kriging in the engine's FP64 order (Cholesky, Schur
scalar, butterfly sums), pinned against the reference's bordered LU
(crates/estimation/src/block.rs:122-198) and numpy: equal to 1e-12 on
well-conditioned systems; on ill-conditioned ones (Gaussian covariance, tiny
nugget, centimetre twins) the gap stays within what the condition number
allows, kappa * eps. CPU only."""
import numpy as np
import pytest

import oracle
from boitata_amd import _abi

EPS = np.finfo(np.float64).eps


def lists_of(n_targets, n):
    return np.tile(np.arange(n, dtype=np.int32), (n_targets, 1)), np.full(n_targets, n, np.int32)


def rel(a, b, floor):
    return np.abs(a - b) / np.maximum(np.abs(b), floor)


def test_engine_order_equals_lu_on_well_conditioned_blocks():
    rng = np.random.default_rng(3)
    n = 40
    P = rng.uniform(-150, 150, (n, 3))
    d = oracle.Data(P[:, 0], P[:, 1], P[:, 2], rng.lognormal(0, 1, n))
    vg = _abi.make_variogram([("exponential", 0.6, 120.0), ("spherical", 0.3, 300.0)], nugget=0.1,
                             rotation=(30, 20, 0), ratios=(0.625, 0.2))
    blk = _abi.make_block((10, 10, 5), (4, 4, 2))
    T = rng.uniform(-100, 100, (25, 3))
    idx, cnt = lists_of(len(T), n)
    lu = oracle.block_krige_lists(d, vg, blk, T, idx, cnt, solver=0)
    ch = oracle.block_krige_lists(d, vg, blk, T, idx, cnt, solver=1, lanes=np.full(len(T), 64))
    assert (lu["status"] == 0).all() and (ch["status"] == 0).all()
    for k in ("value", "variance", "lagrange", "support"):
        assert rel(ch[k], lu[k], 1e-3).max() <= 1e-12, k


@pytest.mark.parametrize("kind,param", [(_abi.SIMPLE, 1.2), (_abi.INDICATOR, 1.0)])
def test_engine_order_kinds(kind, param):
    rng = np.random.default_rng(4)
    n = 20
    P = rng.uniform(-80, 80, (n, 3))
    d = oracle.Data(P[:, 0], P[:, 1], P[:, 2], rng.lognormal(0, 1, n))
    vg = _abi.make_variogram([("spherical", 0.8, 150.0)], nugget=0.2)
    blk = _abi.make_block((8, 8, 4), (2, 2, 2))
    T = rng.uniform(-60, 60, (10, 3))
    idx, cnt = lists_of(len(T), n)
    lu = oracle.block_krige_lists(d, vg, blk, T, idx, cnt, solver=0, kind=kind, param=param)
    ch = oracle.block_krige_lists(d, vg, blk, T, idx, cnt, solver=1, lanes=np.full(len(T), 32), kind=kind,
                                  param=param)
    for k in ("value", "variance"):
        assert rel(ch[k], lu[k], 1e-3).max() <= 1e-12, k


def test_engine_order_matches_numpy_bordered_solve():
    # An independent dense solve of the bordered system (block.rs:140-170):
    # a transposed or dropped term in the Cholesky, the solves or the Schur
    # scalar fails here.
    vg = _abi.make_variogram([("gaussian", 1.0, 100.0)], nugget=1e-2)
    rng = np.random.default_rng(5)
    n = 30
    P = rng.uniform(-80, 80, (n, 3))
    d = oracle.Data(P[:, 0], P[:, 1], P[:, 2], rng.uniform(0, 2, n))
    t = np.array([[3.0, -2.0, 1.0]])
    blk = _abi.make_block((10, 10, 5), (4, 4, 2))
    pts = oracle.block_points(t[0], blk)
    A = np.ones((n + 1, n + 1))
    A[n, n] = 0
    A[:n, :n] = oracle.system_matrix(d, vg, np.arange(n))
    b = np.ones(n + 1)
    b[:n] = [np.mean([oracle.block_cov_points(vg, P[i], q) for q in pts]) for i in range(n)]
    x = np.linalg.solve(A, b)
    cbb = np.mean([[oracle.block_cov_points(vg, p, q) for q in pts] for p in pts])
    idx, cnt = lists_of(1, n)
    e = oracle.block_krige_lists(d, vg, blk, t, idx, cnt, solver=1, lanes=[32])
    assert abs(e["value"][0] - x[:n] @ d.value) < 1e-9
    assert abs(e["lagrange"][0] - x[n]) < 1e-9
    assert abs(e["variance"][0] - max(cbb - x[:n] @ b[:n] - x[n], 0)) < 1e-9


def test_system_matrix_is_cov_points():
    vg = _abi.make_variogram([("spherical", 1.0, 50.0)], nugget=0.5)
    d = oracle.Data([0.0, 10.0], [0.0, 0.0], [0.0, 0.0], [1.0, 2.0], err=[0.25, 0.0])
    M = oracle.system_matrix(d, vg, [0, 1])
    assert M[0, 0] == pytest.approx(1.75) and M[1, 1] == pytest.approx(1.5)
    assert M[0, 1] == M[1, 0] == pytest.approx(oracle.cov(vg, 10.0))


def test_ill_conditioned_gap_is_bounded_by_kappa():
    # Gaussian covariance, nugget 1e-4 and twins 1 cm apart, as the
    # precision-stressed block workload: the two orders differ, by at most a
    # modest multiple of kappa * eps.
    rng = np.random.default_rng(6)
    n = 64
    P = rng.uniform(-60, 60, (n, 3))
    P[1::8] = P[0::8] + rng.normal(0, 0.01, (8, 3))
    d = oracle.Data(P[:, 0], P[:, 1], P[:, 2], rng.lognormal(0, 1, n))
    vg = _abi.make_variogram([("gaussian", 1.0, 100.0)], nugget=1e-4)
    blk = _abi.make_block((10, 10, 5), (4, 4, 2))
    T = rng.uniform(-30, 30, (12, 3))
    idx, cnt = lists_of(len(T), n)
    lu = oracle.block_krige_lists(d, vg, blk, T, idx, cnt, solver=0)
    ch = oracle.block_krige_lists(d, vg, blk, T, idx, cnt, solver=1, lanes=np.full(len(T), 64))
    kappa = np.linalg.cond(oracle.system_matrix(d, vg, idx[0]))
    assert kappa > 1e4  # genuinely ill-conditioned
    assert (ch["status"] == lu["status"]).all()
    for k in ("value", "variance"):
        gap = rel(ch[k], lu[k], 1e-3)
        assert gap.max() <= 1e-12 + 64 * kappa * EPS, (k, gap.max(), kappa)


def test_coincident_and_rejected_systems_take_the_lu():
    # Exact twins: the engine merges them in its LU kernel, so its order
    # defers to block_krige_points (equal results, bit for bit).
    vg = _abi.make_variogram([("spherical", 1.0, 100.0)], nugget=0.1)
    d = oracle.Data([0.0, 0.0, 30.0, -20.0], [0.0, 0.0, 5.0, 40.0], [0.0] * 4, [1.0, 3.0, 2.0, 0.5])
    blk = _abi.make_block((10, 10, 5), (2, 2, 2))
    T = np.array([[5.0, 5.0, 0.0]])
    idx, cnt = lists_of(1, 4)
    lu = oracle.block_krige_lists(d, vg, blk, T, idx, cnt, solver=0)
    ch = oracle.block_krige_lists(d, vg, blk, T, idx, cnt, solver=1, lanes=[16])
    for k in ("value", "variance", "lagrange"):
        assert ch[k][0] == lu[k][0]
