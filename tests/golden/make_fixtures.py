"""Writes the golden fixtures of this directory (run from the repository root:
``python tests/golden/make_fixtures.py``).

Inputs are seeded: the reference's LCG clouds (crates/estimation/src/search.rs:
1184-1203: state * 6364136223846793005 + 1442695040888963407, (state >> 11) /
2^53) and a small drill-hole set of the C1 generator. Outputs come from the
oracle (the FP64 restatement of the reference's path) and are cross-checked
here, before writing, against independent computations:
  * neighbour lists of uncapped searches against a brute-force sort by
    (d2 bits, index) in the search frame, and tree against scan;
  * ordinary kriging against a dense numpy solve of each bordered system;
  * block kriging against a dense numpy solve with numpy-averaged
    point-to-block covariances.
Arrays are stored with numpy.savez_compressed (no pickles)."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import oracle  # noqa: E402
from boitata_amd import _abi, datasets  # noqa: E402
from conftest import Lcg  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def lcg_cloud(seed, n):
    g = Lcg(seed)
    pts = np.array([(g.next() * 1000.0, g.next() * 1000.0, g.next() * 100.0) for _ in range(n)])
    return g, pts, (np.arange(n) // 20).astype(np.int32)


def brute_force(pts, F, t, radius, K):
    d = (pts - t) @ np.asarray(F).reshape(3, 3).T
    d2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
    keys = sorted((int(np.float64(v).view(np.uint64)), i) for i, v in enumerate(d2) if v <= radius * radius)
    return [i for _, i in keys[:K]]


def search_fixture():
    g, pts, holes = lcg_cloud(12345, 3000)
    values = (np.arange(3000) % 10).astype(np.float64)
    T = np.array([(g.next() * 1000.0, g.next() * 1000.0, g.next() * 100.0) for _ in range(300)])
    d = oracle.Data(pts[:, 0], pts[:, 1], pts[:, 2], values, None, holes)
    cases = [
        dict(min_samples=1, max_samples=16, radius=120.0),
        dict(min_samples=1, max_samples=24, radius=200.0, octant=True),
        dict(min_samples=1, max_samples=12, radius=150.0, max_per_hole=3, rotation=(35, 10, 0), ratios=(0.4, 0.2)),
        dict(min_samples=1, max_samples=16, radius=150.0, rotation=(35, 10, 0), ratios=(0.4, 0.2)),
    ]
    out = dict(x=pts[:, 0], y=pts[:, 1], z=pts[:, 2], value=values, hole=holes, targets=T)
    for c, kw in enumerate(cases):
        s = _abi.make_search(**kw)
        idx, cnt, st = oracle.neighbors(d, s, None, T)
        i2, c2, s2 = oracle.neighbors(d, s, None, T, tree=False)
        assert (idx == i2).all() and (cnt == c2).all() and (st == s2).all()
        if "octant" not in kw and "max_per_hole" not in kw:
            F = np.eye(3).ravel()
            if "rotation" in kw:
                F = oracle.aniso_matrix(*kw["rotation"], 1.0, *kw["ratios"])
            for t in range(len(T)):
                assert list(idx[t, :cnt[t]]) == brute_force(pts, F, T[t], kw["radius"], kw["max_samples"]), (c, t)
        out[f"case{c}_idx"] = idx
        out[f"case{c}_count"] = cnt
        out[f"case{c}_status"] = st
        out[f"case{c}_params"] = np.array([kw["min_samples"], kw["max_samples"], kw["radius"],
                                           1.0 if kw.get("octant") else 0.0, kw.get("max_per_hole", -1),
                                           *kw.get("rotation", (0, 0, 0)), *kw.get("ratios", (1.0, 1.0))])
    np.savez_compressed(os.path.join(HERE, "lcg_search.npz"), **out)


def dense_ok(vg, P, t):
    n = len(P)
    A = np.ones((n + 1, n + 1))
    A[n, n] = 0.0
    for i in range(n):
        for j in range(n):
            A[i, j] = oracle.cov_points(vg, P[i], P[j])
    b = np.ones(n + 1)
    b[:n] = [oracle.cov_points(vg, p, t) for p in P]
    return np.linalg.solve(A, b)


def kriging_fixture():
    w = datasets.config1()
    keep = np.arange(0, w.n, 2)
    x, y, z, v, hole = w.x[keep], w.y[keep], w.z[keep], w.value[keep], w.hole[keep]
    T = w.targets[::100].copy()
    d = oracle.Data(x, y, z, v, None, hole)
    vg, s = w.variogram, w.search
    o = oracle.krige(d, vg, s, T)
    idx, cnt, st = oracle.neighbors(d, s, vg, T)
    P = np.stack([x, y, z], 1)
    for t in range(0, len(T), 7):
        if o["status"][t] != 0:
            continue
        sol = dense_ok(vg, P[idx[t, :cnt[t]]], T[t])
        ref = float(sol[:cnt[t]] @ v[idx[t, :cnt[t]]])
        assert abs(o["value"][t] - ref) <= 1e-9 * max(abs(ref), 1e-3), t
    blk = _abi.make_block((10.0, 10.0, 10.0), (2, 2, 2))
    ob = oracle.block_krige(d, vg, s, blk, T)
    pts = oracle.block_points(T[0], blk)
    t0 = 0
    if ob["status"][t0] == 0:
        Q = P[idx[t0, :cnt[t0]]]
        n = len(Q)
        A = np.ones((n + 1, n + 1))
        A[n, n] = 0.0
        for i in range(n):
            for j in range(n):
                A[i, j] = oracle.cov_points(vg, Q[i], Q[j])
        b = np.ones(n + 1)
        b[:n] = [np.mean([oracle.block_cov_points(vg, q, p) for p in pts]) for q in Q]
        sol = np.linalg.solve(A, b)
        ref = float(sol[:n] @ v[idx[t0, :cnt[t0]]])
        assert abs(ob["value"][t0] - ref) <= 1e-9 * max(abs(ref), 1e-3)
    vgp = np.array([vg.nugget, vg.structures[0].sill, vg.structures[0].range, vg.azimuth, vg.dip, vg.rake,
                    vg.ratio_semi, vg.ratio_minor])
    sp = np.array([s.min_samples, s.max_samples, s.radius, s.azimuth, s.dip, s.rake, s.ratio_semi, s.ratio_minor])
    np.savez_compressed(os.path.join(HERE, "c1_kriging.npz"), x=x, y=y, z=z, value=v, hole=hole, targets=T,
                        variogram=vgp, search=sp, idx=idx, count=cnt, status=o["status"], est=o["value"],
                        variance=o["variance"], lagrange=o["lagrange"], block_est=ob["value"],
                        block_variance=ob["variance"], block_status=ob["status"])


if __name__ == "__main__":
    search_fixture()
    kriging_fixture()
    print("fixtures written to", HERE)
