"""The oracle's search pinned by the reference's search tests and by brute
force: every sample within the radius in increasing (d2 bits, index), then
the greedy rules (crates/estimation/src/search.rs:1010-1850,
crates/estimation/src/search/grid.rs:383-538). CPU only."""
import math

import numpy as np

import oracle
from boitata_amd import _abi
from conftest import Lcg


def s_data(pts, holes=None, values=None):
    P = np.asarray(pts, dtype=np.float64)
    if P.shape[1] == 2:
        P = np.c_[P, np.zeros(len(P))]
    v = np.zeros(len(P)) if values is None else values
    return oracle.Data(P[:, 0], P[:, 1], P[:, 2], v, None, holes)


def nb(d, search, targets, vg=None, tree=True, left_out=None):
    idx, count, status = oracle.neighbors(d, search, vg, np.atleast_2d(targets), left_out=left_out, tree=tree)
    return [(list(idx[i, :count[i]]) if status[i] == 0 else None) for i in range(len(count))]


def brute(d, search, target, frame=None):
    """Greedy selection over a numpy sort by (d2 bits, index); no caps."""
    P = np.stack([d.x, d.y, d.z], 1)
    F = np.eye(3) if frame is None else frame
    first = next(i for i in range(d.n) if np.isfinite(P[i]).all())
    o = P[first]
    Q = (P - o) @ F.T
    q = F @ (np.asarray(target) - o)
    d2 = ((Q[:, 0] - q[0]) ** 2 + (Q[:, 1] - q[1]) ** 2) + (Q[:, 2] - q[2]) ** 2
    keep = np.nonzero(d2 <= search.radius ** 2)[0]
    order = keep[np.lexsort((keep, d2[keep].view(np.uint64)))]
    return list(order[: search.max_samples])


def test_selects_nearest_within_radius():
    d = s_data([(1, 0), (2, 0), (100, 0)])
    p = _abi.make_search(1, 2, 10.0)
    assert nb(d, p, (0, 0, 0)) == [[0, 1]]
    assert nb(d, p, (0, 0, 0), tree=False) == [[0, 1]]


def test_respects_max_per_hole():
    d = s_data([(1, 0), (2, 0), (3, 0), (4, 0)], holes=[1, 1, 1, 2])
    p = _abi.make_search(1, 10, 100.0, max_per_hole=2)
    assert nb(d, p, (0, 0, 0), tree=False) == [[0, 1, 3]]
    assert nb(d, p, (0, 0, 0)) == [[0, 1, 3]]


def test_errors_when_too_few():
    d = s_data([(1, 0)])
    p = _abi.make_search(5, 10, 100.0)
    assert nb(d, p, (0, 0, 0)) == [None] and nb(d, p, (0, 0, 0), tree=False) == [None]


def lcg_cloud(seed, n):
    g = Lcg(seed)
    pts, holes = [], []
    for i in range(n):
        pts.append((g.next() * 1000.0, g.next() * 1000.0, g.next() * 100.0))
        holes.append(i // 20)
    return g, pts, holes


def test_tree_matches_the_scan():
    # search.rs:1184-1270 (the cases without domains or high-grade rules).
    g, pts, holes = lcg_cloud(12345, 3000)
    d = s_data(pts, holes=np.asarray(holes, np.int32), values=np.arange(3000) % 10 * 1.0)
    aniso = dict(rotation=(35, 10, 0), ratios=(0.4, 0.2))
    cases = [
        _abi.make_search(1, 16, 120.0),
        _abi.make_search(1, 24, 200.0, octant=True),
        _abi.make_search(1, 12, 150.0, max_per_hole=3, **aniso),
        _abi.make_search(1, 16, 150.0, octant=True),
        _abi.make_search(1, 16, 150.0, **aniso),
        _abi.make_search(1, 24, 60.0, octant=True, max_per_octant=2, min_sectors=6, max_empty_sectors=1,
                         max_per_hole=2, min_holes=6),
        _abi.make_search(1, 16, 150.0, sectors=6, max_per_sector=3, min_sectors=4),
    ]
    for p in cases:
        T = np.array([(g.next() * 1000.0, g.next() * 1000.0, g.next() * 100.0) for _ in range(200)])
        assert nb(d, p, T) == nb(d, p, T, tree=False)


def test_tree_matches_scan_among_ties_far_from_origin():
    x0, y0, z0 = 612_345.0, 7_654_321.0, 850.0
    pts, holes = [], []
    for i in range(9 * 9 * 12):
        a, b, c = i % 9, i // 9 % 9, i // 81
        pts.append((x0 + a * 10.0, y0 + b * 10.0, z0 - c * 2.5))
        holes.append(i % 81)
    d = s_data(pts, holes=np.asarray(holes, np.int32))
    for aniso in [None, (0, 0), (90, 0), (45, 30)]:
        kw = {} if aniso is None else dict(rotation=(aniso[0], aniso[1], 0), ratios=(0.5, 0.25))
        for octant, mph in [(False, None), (True, None), (False, 2), (True, 3)]:
            for radius in (10.0, 20.0, 25.0, 40.0):
                p = _abi.make_search(1, 16, radius, octant=octant, max_per_hole=mph, **kw)
                T = [(x0 + (k % 9) * 5.0, y0 + (k // 9) * 10.0, z0 - (k % 5) * 5.0) for k in range(80)]
                assert nb(d, p, T) == nb(d, p, T, tree=False)


def test_plain_search_is_the_brute_force_order():
    # grid.rs:400-460: the visit order is a sort by (d2 bits, index).
    g, pts, _ = lcg_cloud(7, 3000)
    d = s_data(pts)
    for radius in (50.0, 200.0, math.inf):
        p = _abi.make_search(1, 40, radius)
        T = np.array([(g.next() * 1600.0 - 300, g.next() * 1600.0 - 300, g.next() * 160.0 - 30) for _ in range(100)])
        got = nb(d, p, T)
        for t, gsel in zip(T, got):
            b = brute(d, p, t)
            assert (gsel if gsel is not None else b[: len(b)]) == b


def test_anisotropic_search_is_brute_force_in_the_frame():
    g, pts, _ = lcg_cloud(21, 2000)
    d = s_data(pts)
    p = _abi.make_search(1, 24, 120.0, rotation=(30, 20, 0), ratios=(0.625, 0.2))
    F = oracle.aniso_matrix(30, 20, 0, 1, 0.625, 0.2)
    T = np.array([(g.next() * 1000.0, g.next() * 1000.0, g.next() * 100.0) for _ in range(100)])
    for t, sel in zip(T, nb(d, p, T)):
        b = brute(d, p, t, F)
        if sel is not None:
            assert sel == b


def test_duplicates_come_in_index_order_and_nan_never():
    g, cloud, _ = lcg_cloud(3, 50)
    pts = [(5.0, 5.0, 5.0)] * 100 + [(math.nan, 0.0, 0.0)] + [(x / 100, y / 100, z / 10) for x, y, z in cloud]
    d = s_data(pts)
    p = _abi.make_search(1, 151, math.inf)
    sel = nb(d, p, (5, 5, 5))[0]
    assert 100 not in sel and sel[:100] == list(range(100))


def test_flat_and_duplicate_data():
    pts = [((i % 25) * 3.0, (i // 25) * 3.0) for i in range(500)] + [(10.0, 10.0)] * 100
    d = s_data(pts)
    p = _abi.make_search(1, 12, 20.0)
    for t in [(5, 7, 0), (30, 40, 4), (10, 10, 0)]:
        a, b = nb(d, p, t)[0], nb(d, p, t, tree=False)[0]
        assert a == b


def test_octants_follow_the_ellipse():
    sa, ca = math.sin(math.radians(30)), math.cos(math.radians(30))
    at = lambda along, across, z: (along * sa + across * ca, along * ca - across * sa, z)  # noqa: E731
    pts = [at(along, across, 0.0) for across, along in zip([-0.5, 0.5, -0.4, 0.4], [4.0, 5.0, 6.0, 7.0])]
    pts.append((-20.0, -20.0, 0.0))
    d = s_data(pts)
    p = _abi.make_search(1, 4, 100.0, octant=True, rotation=(30, 0, 0), ratios=(0.5, 0.5))
    assert sorted(nb(d, p, (0, 0, 0), tree=False)[0]) == [0, 1, 4]
    assert sorted(nb(d, p, (0, 0, 0))[0]) == [0, 1, 4]


def test_planar_octants_fill_max_samples():
    d = s_data([(i % 20, i // 20) for i in range(400)])
    for kw in [{}, dict(rotation=(35, 0, 0), ratios=(0.5, 0.5))]:
        p = _abi.make_search(1, 24, 100.0, octant=True, **kw)
        assert len(nb(d, p, (9.3, 10.6, 0))[0]) == 24
        assert len(nb(d, p, (9.3, 10.6, 0), tree=False)[0]) == 24


def test_leave_one_out_is_the_search_without_the_sample():
    # batch.rs:376-460: neighbours around sample i equal those of a tree without i.
    g, pts, holes = lcg_cloud(99, 800)
    d = s_data(pts)
    p = _abi.make_search(1, 16, 150.0, octant=True)
    which = np.arange(0, 800, 7)
    got = nb(d, p, np.asarray(pts)[which], left_out=which)
    for i, sel in zip(which, got):
        keep = np.r_[0:i, i + 1:800]
        rest = s_data(np.asarray(pts)[keep])
        ref = nb(rest, p, pts[i])[0]
        ref = None if ref is None else [int(keep[k]) for k in ref]
        if i != 0:  # the origin of the frame moves when sample 0 leaves
            assert sel == ref


def test_rules_are_checked():
    bad = [
        _abi.make_search(0, 0, 10.0),
        _abi.make_search(5, 4, 10.0),
        _abi.make_search(1, 4, 0.0),
        _abi.make_search(1, 4, 10.0, max_per_octant=2),
        _abi.make_search(1, 4, 10.0, min_sectors=2),
        _abi.make_search(1, 4, 10.0, octant=True, min_sectors=9),
        _abi.make_search(1, 4, 10.0, octant=True, max_empty_sectors=4),
        _abi.make_search(1, 4, 10.0, min_holes=5),
        _abi.make_search(1, 4, 10.0, sectors=1, max_per_sector=2),
        _abi.make_search(1, 4, 10.0, octant=True, sectors=4, max_per_sector=2),
    ]
    for s in bad:
        assert not oracle.check_search(s)
    assert oracle.check_search(_abi.make_search(1, 4, 10.0, octant=True, min_sectors=4, max_empty_sectors=3))
