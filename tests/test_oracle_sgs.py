"""The oracle's sequential Gaussian simulation (crates/simulation/src/sgs.rs
node loop in scores, restated in oracle/oracle.cpp or_sgs) pinned without a
GPU:
- against an independent numpy loop for small cases: brute-force (d2, index)
  order within the radius, simple kriging (mean 0) by numpy.linalg.solve of
  the covariance matrix, score = w . s + sqrt(var) z, the node appended;
- its plan form (or_sgs_plan: one grid over data and nodes, per-node index
  limits, parallel solves, draws in path order) against it, on octants, hole
  caps, planar data under a 3-D grid (the sectors leave the plane with the
  first node off it) and nodes on data;
- the reference's own SGS tests (sgs.rs): a node on a datum takes the datum's
  value (nodes_on_data_take_the_data_value), a hole cap of one on a single
  hole equals max_samples 1 (max_per_hole_caps_the_data_of_one_hole), a node
  with nothing in reach draws from the marginal, and the variogram is scaled to
  a unit sill (unit_sill). The draws and path are inputs: the reference's
  generator (rand's StdRng, shuffle and Normal) is not restated, parity
  unpinned for that stream."""
import numpy as np
import pytest

import oracle
from boitata_amd import _abi


def sph(t):
    t = np.minimum(t, 1.0)
    return 1.5 * t - 0.5 * t ** 3


def numpy_sgs(D, s, G, path, z, nugget, sill, rng_a, K, radius):
    """Isotropic spherical model, no caps: the sequential loop by brute force."""
    c0 = nugget + sill
    nug, sil = nugget / c0, sill / c0  # unit sill

    def cov(h):
        return np.where(h > 0, 1.0 - (nug + sil * sph(h / rng_a)), 1.0)

    P = [tuple(p) for p in D]
    vals = list(s)
    out = np.full(len(G), np.nan)
    src = np.zeros(len(G), np.int32)
    draw = 0
    for node in path:
        t = G[node]
        X = np.array(P)
        d = X - t
        d2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        inside = np.nonzero(d2 <= radius * radius)[0]
        order = inside[np.lexsort((inside, d2[inside]))][:K]
        on = [i for i in order if d2[i] == 0.0]
        if on:
            out[node], src[node] = vals[on[0]], on[0]
            continue
        if len(order) == 0:
            sc = z[draw]
            src[node] = -2
        else:
            S = X[order]
            H = np.sqrt(((S[:, None, :] - S[None, :, :]) ** 2).sum(-1))
            A = cov(H)
            b = cov(np.sqrt(d2[order]))
            w = np.linalg.solve(A, b)
            sc = w @ np.array(vals)[order] + np.sqrt(max(1.0 - w @ b, 0.0)) * z[draw]
            src[node] = -1
        draw += 1
        out[node] = sc
        P.append(tuple(t))
        vals.append(sc)
    return out, src


def data(seed, n, extent=100.0):
    rng = np.random.default_rng(seed)
    D = rng.uniform(0, extent, (n, 3))
    return rng, D, rng.normal(size=n)


def test_oracle_matches_a_numpy_loop():
    rng, D, s = data(1, 40)
    G = rng.uniform(0, 100, (150, 3))
    G[7] = D[11]  # a node on a datum
    G[8] = (500.0, 500.0, 500.0)  # out of reach: the marginal
    path = rng.permutation(len(G))
    z = rng.normal(size=len(G))
    vg = _abi.make_variogram([("spherical", 2.0, 45.0)], nugget=0.4)
    sr = _abi.make_search(1, 10, 60.0)
    got, src = oracle.sgs(oracle.Data(D[:, 0], D[:, 1], D[:, 2], s), vg, sr, G, path, z)
    want, wsrc = numpy_sgs(D, s, G, path, z, 0.4, 2.0, 45.0, 10, 60.0)
    assert (src == wsrc).all()
    assert src[7] == 11 and got[7] == s[11]
    # Node 8's draw: one per node before it on the path that is not on a datum.
    at8, at7 = list(path).index(8), list(path).index(7)
    assert src[8] == -2 and got[8] == z[at8 - (1 if at7 < at8 else 0)]
    assert np.max(np.abs(got - want) / np.maximum(np.abs(want), 1e-3)) <= 1e-10


@pytest.mark.parametrize("kw", [
    dict(),
    dict(octant=True, rotation=(30, 20, 0), ratios=(0.6, 0.4)),
    dict(max_per_hole=2),
    dict(sectors=6, max_per_sector=2),
])
def test_plan_form_equals_the_sequential_loop(kw):
    rng, D, s = data(2, 300)
    hole = (np.arange(300) // 6).astype(np.int32)
    G = rng.uniform(0, 100, (1500, 3))
    G[:20] = D[:20]
    path = rng.permutation(len(G))
    z = rng.normal(size=len(G))
    vg = _abi.make_variogram([("exponential", 0.7, 60.0), ("spherical", 0.3, 30.0)], nugget=0.05)
    sr = _abi.make_search(2, 16, 45.0, **kw)
    d = oracle.Data(D[:, 0], D[:, 1], D[:, 2], s, None, hole)
    a, sa = oracle.sgs(d, vg, sr, G, path, z)
    b, sb = oracle.sgs(d, vg, sr, G, path, z, plan=True, threads=4)
    assert (sa == sb).all()
    assert np.array_equal(a, b)
    assert (sa[:20] == np.arange(20)).all() and np.array_equal(a[:20], s[:20])


def test_planar_data_under_a_solid_grid():
    rng = np.random.default_rng(3)
    D = np.concatenate([rng.uniform(0, 100, (200, 2)), np.zeros((200, 1))], 1)
    s = rng.normal(size=200)
    G = np.concatenate([rng.uniform(0, 100, (600, 2)), rng.choice([0.0, 5.0], (600, 1), p=[0.7, 0.3])], 1)
    path = rng.permutation(len(G))
    z = rng.normal(size=len(G))
    vg = _abi.make_variogram([("spherical", 1.0, 40.0)], nugget=0.1)
    sr = _abi.make_search(1, 12, 50.0, octant=True)
    d = oracle.Data(D[:, 0], D[:, 1], D[:, 2], s)
    a, sa = oracle.sgs(d, vg, sr, G, path, z)
    b, sb = oracle.sgs(d, vg, sr, G, path, z, plan=True)
    assert (sa == sb).all() and np.array_equal(a, b)


def test_reference_properties():
    # nodes_on_data_take_the_data_value (sgs.rs:1113-1141) in scores.
    D = np.array([[0.0, 0, 0], [10, 0, 0], [20, 0, 0]])
    s = np.array([-1.0, 0.5, 1.5])
    G = np.array([[10.0, 0, 0], [5, 0, 0], [15, 0, 0]])
    vg = _abi.make_variogram([("spherical", 1.0, 30.0)])
    sr = _abi.make_search(1, 8, 50.0)
    sc, src = oracle.sgs(oracle.Data(D[:, 0], D[:, 1], D[:, 2], s), vg, sr, G, [0, 1, 2], [0.3, -0.2, 0.9])
    assert src[0] == 1 and sc[0] == 0.5 and (src[1:] == -1).all()
    # max_per_hole_caps_the_data_of_one_hole (sgs.rs:1195-1220): one hole
    # capped at 1 equals max_samples 1.
    locs = np.stack([np.zeros(10), np.zeros(10), np.arange(10.0)], 1)
    vals = np.array([(i * 7 % 10) for i in range(10)], float) / 5 - 1
    G = np.array([[3.0, 0.0, 4.4]])
    vg = _abi.make_variogram([("spherical", 1.0, 20.0)])
    holes = np.zeros(10, np.int32)
    d_h = oracle.Data(locs[:, 0], locs[:, 1], locs[:, 2], vals, None, holes)
    d_n = oracle.Data(locs[:, 0], locs[:, 1], locs[:, 2], vals)
    run = lambda d, K, cap: oracle.sgs(d, vg, _abi.make_search(1, K, 1e9, max_per_hole=cap), G, [0], [0.7])[0]
    assert np.array_equal(run(d_h, 8, 1), run(d_n, 1, None))
    assert not np.array_equal(run(d_h, 8, 1), run(d_n, 8, None))
    assert np.array_equal(run(d_h, 8, None), run(d_n, 8, None))
    # Nothing in reach: the node is the draw.
    sc, src = oracle.sgs(d_n, vg, _abi.make_search(1, 8, 5.0), np.array([[100.0, 100, 100]]), [0], [1.25])
    assert src[0] == -2 and sc[0] == 1.25
    # unit_sill (sgs.rs:881-886): the model's sills only matter relative to each other.
    rng, D, s = data(4, 60)
    G = rng.uniform(0, 100, (80, 3))
    path, z = rng.permutation(80), rng.normal(size=80)
    d = oracle.Data(D[:, 0], D[:, 1], D[:, 2], s)
    a = oracle.sgs(d, _abi.make_variogram([("spherical", 1.0, 40.0)], nugget=0.25), sr, G, path, z)[0]
    b = oracle.sgs(d, _abi.make_variogram([("spherical", 4.0, 40.0)], nugget=1.0), sr, G, path, z)[0]
    assert np.max(np.abs(a - b)) <= 1e-12
