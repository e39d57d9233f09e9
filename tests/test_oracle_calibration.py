"""The oracle's calibrated search (kriging_algebra.rs:320-441, search.rs:909-936)
pinned without the reference: prefix qualities against dense solves of each
leading system and against the oracle's own kriging of it (as the reference's
prefix_quality_matches_direct_solves does), calibrated_count's rules, the
driver's trimmed counts, and the balanced order of an octant search
(batch.rs:736-770). Samples follow the reference tests' generator
(kriging_algebra.rs:1072-1097)."""
import math

import numpy as np
import pytest

import oracle
from boitata_amd import _abi

from conftest import Lcg


def nested():
    return _abi.make_variogram([("spherical", 0.6, 40.0), ("exponential", 0.3, 120.0)], nugget=0.1)


def scattered(g, n):
    """n samples around the origin, nearest first, every fifth with error variance 0.05."""
    pts, vals, err = [], [], []
    for i in range(n):
        x, y, z = g.next() * 100.0 - 50.0, g.next() * 100.0 - 50.0, g.next() * 20.0
        pts.append((x, y, z))
        vals.append(g.next() * 3.0)
        err.append(0.05 if i % 5 == 0 else 0.0)
    pts = np.array(pts)
    d = np.hypot(np.hypot(pts[:, 0], pts[:, 1]), pts[:, 2])
    o = np.argsort(d, kind="stable")
    return pts[o], np.array(vals)[o], np.array(err)[o]


def close(a, b):
    return abs(a - b) <= 1e-10 * (1.0 + max(abs(a), abs(b)))


def dense_quality(vg, P, err, rhs, support, ordinary):
    """slope, efficiency of kriging from all of P with a dense FP64 solve."""
    n = len(P)
    C = np.array([[oracle.cov_points(vg, P[i], P[j]) for j in range(n)] for i in range(n)]) + np.diag(err)
    a = np.linalg.solve(C, rhs)
    if not ordinary:
        cov, mu = float(a @ rhs), 0.0
    else:
        b = np.linalg.solve(C, np.ones(n))
        mu = (a.sum() - 1.0) / b.sum()
        lam = a - mu * b
        cov = float(lam @ rhs)
    return cov / (cov - mu), (cov + mu) / support


def test_prefix_quality_matches_direct_solves():
    vg = nested()
    t = np.array([1.0, 2.0, 5.0])
    P, v, err = scattered(Lcg(17), 24)
    d = oracle.Data(P[:, 0], P[:, 1], P[:, 2], v, err)
    idx = np.arange(24)
    sill = oracle.total_sill(vg)
    rhs = np.array([oracle.cov_points(vg, P[i], t) for i in range(24)])
    ok_s, ok_e = oracle.prefix_quality(d, vg, t, idx, _abi.ORDINARY)
    sk_s, sk_e = oracle.prefix_quality(d, vg, t, idx, _abi.SIMPLE)
    blk = _abi.make_block((10.0, 10.0, 5.0), (3, 3, 2))
    pts = oracle.block_points(t, blk)
    brhs = np.array([np.mean([oracle.block_cov_points(vg, P[i], q) for q in pts]) for i in range(24)])
    cbb = np.mean([[oracle.block_cov_points(vg, p, q) for q in pts] for p in pts])
    bk_s, bk_e = oracle.prefix_quality(d, vg, t, idx, _abi.ORDINARY, block=blk)
    for n in range(1, 25):
        sub = oracle.Data(P[:n, 0], P[:n, 1], P[:n, 2], v[:n], err[:n])
        # Dense solves of the leading system.
        s, e = dense_quality(vg, P[:n], err[:n], rhs[:n], sill, True)
        assert close(ok_s[n - 1], s) and close(ok_e[n - 1], e), n
        s, e = dense_quality(vg, P[:n], err[:n], rhs[:n], sill, False)
        assert close(sk_s[n - 1], 1.0) and close(sk_e[n - 1], e), n
        s, e = dense_quality(vg, P[:n], err[:n], brhs[:n], cbb, True)
        assert close(bk_s[n - 1], s) and close(bk_e[n - 1], e), n
        # The oracle's own kriging of the leading samples (Estimate::slope and
        # efficiency, krige.rs:44-66).
        k = oracle.krige_system(sub, vg, t)
        var, mu = k["variance"], k["lagrange"]
        assert k["status"] == 0
        assert close(ok_e[n - 1], (sill - var) / sill)
        assert close(ok_s[n - 1], (sill - var - mu) / (sill - var - 2.0 * mu))


def test_calibrated_count_takes_the_fewest_that_meet_the_target():
    vg = nested()
    g = Lcg(5)
    for trial in range(40):
        P, v, err = scattered(g, 20)
        t = np.array([g.next() * 20.0, g.next() * 20.0, 10.0])
        d = oracle.Data(P[:, 0], P[:, 1], P[:, 2], v, err)
        slope, eff = oracle.prefix_quality(d, vg, t, np.arange(20))
        q = {_abi.CALIBRATE_SLOPE: slope, _abi.CALIBRATE_EFFICIENCY: eff}
        lo = 1 + trial % 4
        for cal, target in ((_abi.CALIBRATE_SLOPE, 0.8), (_abi.CALIBRATE_SLOPE, 0.95),
                            (_abi.CALIBRATE_EFFICIENCY, 0.3), (_abi.CALIBRATE_EFFICIENCY, 0.99)):
            n, met = oracle.calibrated_count(cal, target, lo, slope, eff)
            assert lo <= n <= 20
            meets = q[cal] >= target
            if met:
                assert meets[n - 1] and not meets[lo - 1:n - 1].any()
            else:
                assert n == 20 and not meets[lo - 1:].any()


def test_driver_keeps_the_calibrated_count():
    vg = nested()
    g = Lcg(29)
    P, v, err = scattered(g, 40)
    d = oracle.Data(P[:, 0], P[:, 1], P[:, 2], v, err)
    T = np.array([[g.next() * 40.0 - 20.0, g.next() * 40.0 - 20.0, g.next() * 10.0] for _ in range(30)])
    trimmed = 0
    for kw in (dict(target_slope=0.9), dict(target_efficiency=0.5), dict(target_efficiency=0.9)):
        s = _abi.make_search(min_samples=3, max_samples=24, radius=200.0, **kw)
        plain = _abi.make_search(min_samples=3, max_samples=24, radius=200.0)
        idx, cnt, st = oracle.neighbors(d, plain, vg, T)
        out = oracle.krige(d, vg, s, T)
        for i in range(len(T)):
            sl, ef = oracle.prefix_quality(d, vg, T[i], idx[i, :cnt[i]])
            n, _ = oracle.calibrated_count(s.calibration, s.calibration_target, 3, sl, ef)
            assert out["n_used"][i] == n
            ref = oracle.krige(oracle.Data(P[idx[i, :n], 0], P[idx[i, :n], 1], P[idx[i, :n], 2], v[idx[i, :n]],
                                           err[idx[i, :n]]), vg, plain, T[i:i + 1])
            assert ref["n_used"][0] == n and ref["value"][0] == out["value"][i]
            trimmed += n < cnt[i]
    assert trimmed > 0


def test_simple_kriging_takes_no_slope_target_and_targets_are_checked():
    vg = nested()
    P, v, err = scattered(Lcg(3), 12)
    d = oracle.Data(P[:, 0], P[:, 1], P[:, 2], v, err)
    T = np.zeros((1, 3))
    with pytest.raises(ValueError):
        oracle.krige(d, vg, _abi.make_search(1, 12, 100.0, target_slope=0.9), T, kind=_abi.SIMPLE, param=1.0)
    out = oracle.krige(d, vg, _abi.make_search(1, 12, 100.0, target_efficiency=0.1), T, kind=_abi.SIMPLE, param=1.0)
    assert out["status"][0] == 0
    for kw in (dict(target_slope=0.0), dict(target_slope=math.nan), dict(target_efficiency=1.5),
               dict(target_efficiency=math.inf)):
        assert not oracle.check_search(_abi.make_search(1, 12, 100.0, **kw))


def test_calibrated_octant_searches_take_the_sectors_in_turn():
    # batch.rs:736-770: 40 samples on a spiral in one plane (quadrants).
    i = np.arange(40)
    r, a = 1.0 + i, i * 0.37
    P = np.stack([r * np.cos(a), r * np.sin(a), np.zeros(40)], 1)
    d = oracle.Data(P[:, 0], P[:, 1], P[:, 2], np.zeros(40))
    t = np.zeros(3)
    octant = _abi.make_search(min_samples=1, max_samples=16, octant=True)
    calibrated = _abi.make_search(min_samples=1, max_samples=16, octant=True, target_slope=0.9)
    idx, cnt, _ = oracle.neighbors(d, octant, None, t[None, :])
    chosen = idx[0, :cnt[0]]
    assert (oracle.balanced(d, octant, None, t, chosen) == chosen).all()
    turns = oracle.balanced(d, calibrated, None, t, chosen)
    assert sorted(turns) == sorted(chosen)
    quadrant = lambda k: (int(P[k, 0] >= 0.0) << 1) | int(P[k, 1] >= 0.0)  # noqa: E731
    assert sorted(quadrant(k) for k in turns[:4]) == [0, 1, 2, 3]
