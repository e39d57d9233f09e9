"""The oracle pinned against the reference's own tests (restated on the same
inputs) and against what the mathematics fixes: closed forms, independent
dense solves, brute force. CPU only."""
import math

import numpy as np
import pytest

import oracle
from boitata_amd import _abi
from conftest import Lcg

close9 = lambda a, b: abs(a - b) <= 1e-9 * (1.0 + max(abs(a), abs(b)))  # noqa: E731  kernel.rs:178-180


# --- crates/core/src/rotation.rs:55-104 ---------------------------------------
def test_azimuth_is_clockwise_from_north():
    assert np.allclose(oracle.rotation_matrix(0, 0, 0)[0], [0, 1, 0], atol=1e-12)
    assert np.allclose(oracle.rotation_matrix(90, 0, 0)[0], [1, 0, 0], atol=1e-12)


def test_dip_is_positive_down():
    assert np.allclose(oracle.rotation_matrix(0, 30, 0)[0], [0, math.cos(math.radians(30)), -0.5], atol=1e-12)


def test_rotations_are_proper():
    for a, d, r in [(0, 0, 0), (37, -12, 81), (300, 60, 15)]:
        m = oracle.rotation_matrix(a, d, r)
        assert np.linalg.norm(m @ m.T - np.eye(3)) < 1e-12
        assert abs(np.linalg.det(m) - 1) < 1e-12
        assert abs(np.linalg.det(oracle.block_frame([a, d, r])) - 1) < 1e-12


def test_block_frame_identity_unrotated_and_y_to_azimuth():
    assert np.linalg.norm(oracle.block_frame([0, 0, 0]) - np.eye(3)) < 1e-12
    assert np.allclose(oracle.block_frame([90, 0, 0])[1], [1, 0, 0], atol=1e-12)


def test_rotation_matches_independent_euler_composition():
    # The three elementary rotations composed with numpy: a transposed operand
    # or a swapped sign fails here.
    for a, d, r in [(30, 20, 0), (137, -40, 65)]:
        sa, ca = math.sin(math.radians(a)), math.cos(math.radians(a))
        sd, cd = math.sin(math.radians(d)), math.cos(math.radians(d))
        sr, cr = math.sin(math.radians(r)), math.cos(math.radians(r))
        r1 = np.array([[sa, ca, 0], [-ca, sa, 0], [0, 0, 1]])
        r2 = np.array([[cd, 0, -sd], [0, 1, 0], [sd, 0, cd]])
        r3 = np.array([[1, 0, 0], [0, cr, sr], [0, -sr, cr]])
        assert np.allclose(oracle.rotation_matrix(a, d, r), r3 @ r2 @ r1, atol=1e-15)


# --- crates/variogram/src/aniso.rs:126-165 ------------------------------------
def test_isotropic_lag_equals_euclidean_scaled():
    m = oracle.aniso_matrix(0, 0, 0, 100, 100, 100)
    assert abs(np.linalg.norm(m @ [30, 40, 0]) - 0.5) < 1e-12


def test_major_axis_matches_direction():
    for az, dip in [(0, 0), (90, 0), (135, 30), (300, -45)]:
        major = oracle.rotation_matrix(az, dip, 17)[0]
        a, d = math.radians(az), math.radians(dip)
        expect = [math.sin(a) * math.cos(d), math.cos(a) * math.cos(d), -math.sin(d)]
        assert np.linalg.norm(major - expect) < 1e-12


def test_variogram_lag_is_frame_norm():
    vg = _abi.make_variogram([("spherical", 1.0, 1.0)], rotation=(40, 10, 0), ratios=(0.5, 0.25))
    m = oracle.aniso_matrix(40, 10, 0, 1, 0.5, 0.25)
    p, q = np.array([1.0, 2, 3]), np.array([20.0, -5, 9])
    h = np.linalg.norm(m @ (q - p))
    assert abs(oracle.cov_points(vg, p, q) - oracle.cov(vg, h)) < 1e-12


# --- crates/variogram/src/model.rs:336-401 ------------------------------------
def test_spherical_bounds():
    assert oracle.shape("spherical", 0, 100) == 0
    assert abs(oracle.shape("spherical", 100, 100) - 1) < 1e-12
    assert oracle.shape("spherical", 200, 100) == 1
    mid = oracle.shape("spherical", 50, 100)
    assert 0 < mid < 1


def test_cubic_pentaspherical_circular():
    assert abs(oracle.shape("cubic", 100, 100) - 1) < 1e-9
    assert oracle.shape("cubic", 150, 100) == 1
    a, b = oracle.shape("pentaspherical", 25, 100), oracle.shape("pentaspherical", 75, 100)
    assert a < b < 1
    v = oracle.shape("circular", 50, 100)
    assert 0 < v < 1 and abs(oracle.shape("circular", 100, 100) - 1) < 1e-9


def test_matern_half_is_exponential_and_orders():
    for h in (10, 55, 130):
        assert abs(oracle.shape("matern", h, 100, 0.5) - oracle.shape("exponential", h, 100)) < 1e-12
    assert oracle.shape("matern", 20, 100, 2.5) < oracle.shape("matern", 20, 100, 0.5)


def test_power_and_sinehole():
    assert abs(oracle.shape("power", 100, 100, 1.5) - 1) < 1e-12
    assert oracle.shape("power", 200, 100, 1.5) > 1
    assert max(oracle.shape("sinehole", 10.0 * i, 100) for i in range(1, 40)) > 1


@pytest.mark.parametrize("model", ["spherical", "exponential", "gaussian", "cubic", "pentaspherical", "circular",
                                   "sinehole"])
def test_shapes_match_closed_forms(model):
    # model.rs:84-125, written independently in numpy.
    r = 80.0
    for h in [0.5, 7.0, 23.0, 61.0, 79.9, 80.0, 140.0]:
        t = h / r
        f = {
            "spherical": 1.0 if t >= 1 else 1.5 * t - 0.5 * t ** 3,
            "exponential": 1 - math.exp(-3 * t),
            "gaussian": 1 - math.exp(-3 * t * t),
            "cubic": 1.0 if t >= 1 else 7 * t ** 2 - 8.75 * t ** 3 + 3.5 * t ** 5 - 0.75 * t ** 7,
            "pentaspherical": 1.0 if t >= 1 else 1.875 * t - 1.25 * t ** 3 + 0.375 * t ** 5,
            "circular": 1.0 if t >= 1 else 1 - 2 / math.pi * math.acos(t) + 2 * t / math.pi * math.sqrt(1 - t * t),
            "sinehole": 1 - math.sin(math.pi * t) / (math.pi * t),
        }[model]
        assert abs(oracle.shape(model, h, r) - f) < 1e-13
    assert oracle.shape(model, 0.0, r) == 0.0
    assert oracle.shape(model, float("nan"), r) == 0.0


def test_derivatives_match_finite_differences():
    # model.rs:356-401 (the three models the adjoint differentiates).
    for model in ("spherical", "exponential", "gaussian"):
        for h in (7.0, 23.0, 61.0, 140.0):
            eps = 1e-4
            d1 = (oracle.shape(model, h + eps, 100) - oracle.shape(model, h - eps, 100)) / (2 * eps)
            assert abs(d1 - oracle.d_shape(model, h / 100)[0] / 100) < 1e-6


# --- crates/variogram/src/composite.rs:192-257 --------------------------------
def test_nugget_discontinuity_and_sill():
    v = _abi.make_variogram([("spherical", 0.8, 100)], nugget=0.2)
    assert oracle.gamma(v, 0.0) == 0.0
    assert abs(oracle.gamma(v, 1e-9) - 0.2) < 0.01
    assert abs(oracle.total_sill(v) - 1) < 1e-12
    assert abs(oracle.gamma(v, 1000.0) - 1) < 1e-9


def test_covariance_complements_semivariance():
    v = _abi.make_variogram([("exponential", 1.0, 100)])
    for h in (5, 40, 90, 300):
        assert abs(oracle.cov(v, h) - (oracle.total_sill(v) - oracle.gamma(v, h))) < 1e-12
    assert abs(oracle.cov(v, 0.0) - oracle.total_sill(v)) < 1e-12


def test_nested_structures_sum():
    v = _abi.make_variogram([("spherical", 0.5, 100), ("exponential", 0.4, 400)], nugget=0.1)
    assert abs(oracle.total_sill(v) - 1) < 1e-12
    expect = 0.1 + 0.5 * oracle.shape("spherical", 50, 100) + 0.4 * oracle.shape("exponential", 50, 400)
    assert abs(oracle.gamma(v, 50.0) - expect) < 1e-15


def test_block_cov_points_leaves_out_the_nugget():
    v = _abi.make_variogram([("spherical", 0.7, 100)], nugget=0.3)
    assert oracle.block_cov_points(v, [0, 0, 0], [0, 0, 0]) == pytest.approx(0.7, abs=1e-15)
    assert oracle.block_cov_points(v, [0, 0, 0], [10, 0, 0]) == oracle.cov(v, 10.0)


def test_per_structure_ratios_reduce_to_shared():
    shared = _abi.make_variogram([("exponential", 0.5, 150.0), ("spherical", 0.4, 400.0)], nugget=0.1,
                                 rotation=(30, 20, 0), ratios=(0.625, 0.2))
    own = _abi.make_variogram([("exponential", 0.5, 150.0, None, (0.625, 0.2)),
                               ("spherical", 0.4, 400.0, None, (0.625, 0.2))], nugget=0.1, rotation=(30, 20, 0),
                              ratios=(0.625, 0.2))
    rng = np.random.default_rng(0)
    for _ in range(50):
        p, q = rng.uniform(-300, 300, 3), rng.uniform(-300, 300, 3)
        assert oracle.cov_points(shared, p, q) == oracle.cov_points(own, p, q)


def test_per_structure_ratios_follow_each_ellipsoid():
    own = _abi.make_variogram([("exponential", 0.5, 150.0, None, (100 / 150, 40 / 150)),
                               ("spherical", 0.4, 400.0, None, (250 / 400, 80 / 400))], nugget=0.1,
                              rotation=(30, 20, 0), ratios=(0.625, 0.2))
    R = oracle.rotation_matrix(30, 20, 0)
    rng = np.random.default_rng(1)
    for _ in range(50):
        p, q = rng.uniform(-300, 300, 3), rng.uniform(-300, 300, 3)
        r = R @ (q - p)
        h1 = np.linalg.norm(r / [150, 100, 40])
        h2 = np.linalg.norm(r / [400, 250, 80])
        g = 0.1 + 0.5 * (1 - math.exp(-3 * h1)) + 0.4 * (1.0 if h2 >= 1 else 1.5 * h2 - 0.5 * h2 ** 3)
        assert abs(oracle.cov_points(own, p, q) - (1.0 - g)) < 1e-12


# --- crates/estimation/src/kernel.rs:125-278 ----------------------------------
def kernel_variogram(nugget=0.1):
    return _abi.make_variogram([("spherical", 0.6, 80.0), ("exponential", 0.3, 200.0)], nugget=nugget,
                               rotation=(30, 10, 0), ratios=(0.5, 0.25))


def kernel_system(g, n):
    o = (650_000.0, 7_400_000.0, 1_200.0)
    xs, ys, zs, vs = [], [], [], []
    for _ in range(n):
        xs.append(o[0] + g.next() * 200.0)
        ys.append(o[1] + g.next() * 200.0)
        zs.append(o[2] + g.next() * 50.0)
        vs.append(g.next() * 3.0)
    return (o[0] + 100.0, o[1] + 100.0, o[2] + 25.0), oracle.Data(xs, ys, zs, vs)


def test_kernel_matches_lu():
    vg = kernel_variogram()
    g = Lcg(7)
    for trial in range(300):
        n = 1 + trial % 40
        target, d = kernel_system(g, n)
        for kind, param in [(_abi.ORDINARY, 0.0), (_abi.SIMPLE, 1.2), (_abi.INDICATOR, 1.5)]:
            fast = oracle.krige_system(d, vg, target, kind, param, solver=1)
            slow = oracle.krige_system(d, vg, target, kind, param, solver=2)
            assert fast["status"] == 0 and slow["status"] == 0
            for k in ("value", "variance", "lagrange"):
                assert close9(fast[k], slow[k]), (trial, kind, k)
            for a, b in zip(fast["weights"], slow["weights"]):
                assert close9(a, b)


def test_kernel_matches_numpy_solve_of_the_bordered_system():
    # An independent dense FP64 solve of the same system (numpy / LAPACK).
    vg = kernel_variogram()
    g = Lcg(17)
    for n in (1, 5, 17, 40, 64):
        target, d = kernel_system(g, n)
        pts = np.stack([d.x, d.y, d.z], 1)
        A = np.ones((n + 1, n + 1))
        A[n, n] = 0.0
        for i in range(n):
            for j in range(n):
                A[i, j] = oracle.cov_points(vg, pts[i], pts[j])
        b = np.ones(n + 1)
        b[:n] = [oracle.cov_points(vg, p, target) for p in pts]
        x = np.linalg.solve(A, b)
        e = oracle.krige_system(d, vg, target, solver=1)
        assert np.allclose(e["weights"], x[:n], rtol=1e-8, atol=1e-10)
        assert abs(e["lagrange"] - x[n]) < 1e-9
        assert abs(e["value"] - x[:n] @ d.value) < 1e-9
        assert abs(e["variance"] - max(oracle.total_sill(vg) - x[:n] @ b[:n] - x[n], 0)) < 1e-9


def test_ordinary_weights_sum_to_one():
    vg = kernel_variogram()
    g = Lcg(11)
    for n in (1, 5, 32, 64):
        target, d = kernel_system(g, n)
        e = oracle.krige_system(d, vg, target)
        assert abs(e["weights"].sum() - 1) < 1e-10


def test_exact_at_a_datum_without_nugget():
    vg = kernel_variogram(0.0)
    g = Lcg(3)
    _, d = kernel_system(g, 20)
    e = oracle.krige_system(d, vg, (d.x[4], d.y[4], d.z[4]), solver=1)
    assert abs(e["value"] - d.value[4]) < 1e-9 and e["variance"] < 1e-9


def test_duplicates_are_singular_before_the_guard():
    vg = kernel_variogram(0.0)
    g = Lcg(5)
    target, d = kernel_system(g, 6)
    d = oracle.Data(np.r_[d.x, d.x[0]], np.r_[d.y, d.y[0]], np.r_[d.z, d.z[0]], np.r_[d.value, d.value[0]])
    assert oracle.krige_system(d, vg, target, solver=1)["status"] == _abi.SINGULAR
    e = oracle.krige_system(d, vg, target)
    assert e["status"] == 0 and e["weights"][0] == e["weights"][6]


def test_falls_back_beyond_max_and_without_a_sill():
    g = Lcg(9)
    target, d = kernel_system(g, 65)
    assert oracle.krige_system(d, kernel_variogram(), target, solver=1)["status"] == -1
    power = _abi.make_variogram([("power", 1.0, 100.0, 1.5)])
    d8 = oracle.Data(d.x[:8], d.y[:8], d.z[:8], d.value[:8])
    assert oracle.krige_system(d8, power, target, solver=1)["status"] == -1
    assert oracle.krige_system(oracle.Data([], [], [], []), kernel_variogram(), target)["status"] == \
        _abi.INSUFFICIENT_DATA


def test_measurement_error_enters_the_diagonal():
    vg = kernel_variogram()
    g = Lcg(13)
    target, d = kernel_system(g, 10)
    err = np.zeros(10)
    err[2] = 0.4
    d = oracle.Data(d.x, d.y, d.z, d.value, err)
    fast = oracle.krige_system(d, vg, target, solver=1)
    slow = oracle.krige_system(d, vg, target, solver=2)
    assert close9(fast["value"], slow["value"]) and close9(fast["variance"], slow["variance"])


# --- crates/estimation/src/krige.rs:219-388 -----------------------------------
def samp(pts, vals):
    pts = np.asarray(pts, dtype=np.float64)
    return oracle.Data(pts[:, 0], pts[:, 1], np.zeros(len(pts)), vals)


def test_exact_interpolation_and_sk_far_away():
    vg = _abi.make_variogram([("spherical", 1.0, 100.0)])
    d = samp([(0, 0), (50, 0), (0, 50)], [5, 2, 3])
    e = oracle.krige_system(d, vg, (0, 0, 0))
    assert abs(e["value"] - 5) < 1e-6 and e["variance"] < 1e-6
    d2 = samp([(0, 0), (50, 0)], [5, 2])
    e = oracle.krige_system(d2, vg, (10000, 10000, 0), _abi.SIMPLE, 3.0)
    assert abs(e["value"] - 3) < 1e-6 and abs(e["variance"] - 1) < 1e-6


def test_efficiency_and_slope():
    vg = _abi.make_variogram([("spherical", 1.0, 100.0)])
    pts = [(0, 0), (50, 0), (0, 50), (60, 70)]
    d = samp(pts, [5, 2, 3, 1])
    t = (90.0, 20.0, 0.0)
    ok = oracle.krige_system(d, vg, t)
    w = ok["weights"]
    P = np.c_[np.asarray(pts, float), np.zeros(4)]
    cov = sum(w[i] * oracle.cov_points(vg, P[i], t) for i in range(4))
    var = sum(w[i] * w[j] * oracle.cov_points(vg, P[i], P[j]) for i in range(4) for j in range(4))
    c0, s2, mu = 1.0, ok["variance"], ok["lagrange"]
    slope = (c0 - s2 - mu) / (c0 - s2 - 2 * mu)  # krige.rs:54-67
    assert abs(slope - cov / var) < 1e-9 and abs((c0 - s2 - 2 * mu) - var) < 1e-9


def test_measurement_error_blends_the_datum():
    vg = _abi.make_variogram([("spherical", 1.0, 100.0)])
    others = [(50, 0), (0, 50), (60, 70)]
    ov = [2.0, 3.0, 1.0]
    for kind, param in [(_abi.ORDINARY, 0.0), (_abi.SIMPLE, 2.0)]:
        loo = oracle.krige_system(samp(others, ov), vg, (0, 0, 0), kind, param)
        for e in (0.1, 0.5, 2.0):
            d = samp([(0, 0)] + others, [5.0] + ov)
            d = oracle.Data(d.x, d.y, d.z, d.value, np.array([e, 0, 0, 0.0]))
            est = oracle.krige_system(d, vg, (0, 0, 0), kind, param)
            k = loo["variance"] / (loo["variance"] + e)
            assert abs(est["value"] - (loo["value"] + k * (5 - loo["value"]))) < 1e-9
            assert abs(est["variance"] - loo["variance"] * (1 - k)) < 1e-9


def test_negative_weights_and_indicator_range():
    vg = _abi.make_variogram([("gaussian", 1.0, 60.0)])
    d = samp([(10, 0), (20, 0), (0, 15), (-30, -5)], [1, 2, 3, 4])
    e = oracle.krige_system(d, vg, (0, 0, 0))
    assert (e["weights"] < 0).any()
    vs = _abi.make_variogram([("spherical", 1.0, 100.0)])
    e = oracle.krige_system(samp([(0, 0), (50, 0), (0, 50)], [1, 4, 2]), vs, (10, 10, 0), _abi.INDICATOR, 2.5)
    assert -1e-6 <= e["value"] <= 1 + 1e-6


# --- crates/estimation/src/coincident.rs:114-257 ------------------------------
def pair_and_lattice(n, gap):
    pts = [(10.0, 10.0), (10.0 + gap, 10.0)]
    vals = [1.0, 5.0]
    k = 0
    while len(pts) < n:
        x, y = (k % 9) * 7.0 - 20.0, (k // 9) * 7.0 - 20.0
        if abs(x - 10) + abs(y - 10) > 1:
            pts.append((x, y))
            vals.append(3.0 + math.sin(k * 0.3))
        k += 1
    return samp(pts, vals)


def test_near_coincident_pair_outcome_independent_of_n():
    vg = _abi.make_variogram([("spherical", 1.0, 50.0)])
    t = (12.0, 11.0, 0.0)
    for gap in (0.0, 1e-12, 1e-7, 1e-6):
        for kind, param in [(_abi.ORDINARY, 0.0), (_abi.SIMPLE, 3.0), (_abi.INDICATOR, 2.0)]:
            a = oracle.krige_system(pair_and_lattice(64, gap), vg, t, kind, param)
            b = oracle.krige_system(pair_and_lattice(65, gap), vg, t, kind, param)
            for e in (a, b):
                assert e["status"] == 0 and e["weights"][0] == e["weights"][1]
            if kind != _abi.SIMPLE:
                assert abs(a["weights"].sum() - 1) < 1e-12
            assert (a["n_used"], b["n_used"]) == (64, 65)


def test_a_repeated_datum_krige_as_its_mean():
    vg = _abi.make_variogram([("exponential", 0.8, 40.0)], nugget=0.2)
    base = pair_and_lattice(20, 0.0)
    ox, oy, ov = base.x[2:], base.y[2:], base.value[2:]
    t = (12.0, 11.0, 0.0)
    three = oracle.Data(np.r_[[10.0] * 3, ox], np.r_[[10.0] * 3, oy], np.zeros(21), np.r_[[1.0, 2.0, 6.0], ov],
                        np.r_[[0.3] * 3, np.zeros(18)])
    one = oracle.Data(np.r_[10.0, ox], np.r_[10.0, oy], np.zeros(19), np.r_[3.0, ov], np.r_[0.1, np.zeros(18)])
    a, b = oracle.krige_system(three, vg, t), oracle.krige_system(one, vg, t)
    assert close9(a["value"], b["value"]) and close9(a["variance"], b["variance"])
    assert close9(3 * a["weights"][0], b["weights"][0]) and close9(a["lagrange"], b["lagrange"])


def test_twins_krige_as_one_sample():
    # crates/estimation/src/lib.rs:112-137
    locs = [(0, 0, 0), (10, 0, 0), (0, 0, 0), (5, 5, -0.0), (5, 5, 0)]
    P = np.asarray(locs, float)
    d = oracle.Data(P[:, 0], P[:, 1], P[:, 2], [1, 2, 3, 4, 6])
    vg = _abi.make_variogram([("spherical", 1.0, 50.0)])
    est = oracle.krige_system(d, vg, (2, 2, 0))
    kept = oracle.Data([0, 10, 5], [0, 0, 5], [0, 0, 0], [2, 2, 5])
    one = oracle.krige_system(kept, vg, (2, 2, 0))
    assert abs(est["value"] - one["value"]) < 1e-12 and abs(est["variance"] - one["variance"]) < 1e-12
    assert est["weights"][0] == est["weights"][2] and est["weights"][3] == est["weights"][4]


def test_indicator_twins_average_their_indicators():
    vg = _abi.make_variogram([("spherical", 1.0, 50.0)])
    d = pair_and_lattice(10, 0.0)
    e = oracle.krige_system(d, vg, (d.x[0], d.y[0], 0.0), _abi.INDICATOR, 2.0)
    assert close9(e["value"], 0.5) and e["variance"] < 1e-9


# --- crates/estimation/src/block.rs:200-370 -----------------------------------
def test_discretization_centres_on_block():
    blk = _abi.make_block((10, 10, 10), (3, 3, 3))
    pts = oracle.block_points([100, 200, 300], blk)
    assert len(pts) == 27 and abs(pts[:, 0].mean() - 100) < 1e-9


def test_block_variance_below_point_variance():
    vg = _abi.make_variogram([("spherical", 1.0, 100.0)])
    d = samp([(0, 0), (60, 0), (0, 60)], [1, 2, 3])
    blk = _abi.make_block((20, 20, 20), (3, 3, 3))
    b = oracle.block_krige_system(d, vg, oracle.block_points([30, 30, 0], blk))
    p = oracle.krige_system(d, vg, (30, 30, 0))
    assert b["variance"] <= p["variance"] + 1e-9 and abs(b["weights"].sum() - 1) < 1e-9


def test_nugget_averages_out_of_the_block():
    c0, c, a, l = 0.4, 1.0, 100.0, 60.0
    blk = _abi.make_block((l, 0, 0), (400, 1, 1))
    d = samp([(-50, 0), (70, 0)], [1, 2])
    vg = _abi.make_variogram([("spherical", c, a)], nugget=c0)
    e = oracle.block_krige_system(d, vg, oracle.block_points([0, 0, 0], blk))
    gamma_bar = c * (l / (2 * a) - l ** 3 / (20 * a ** 3))
    assert abs(e["support"] - (c - gamma_bar)) < 1e-4
    pure = _abi.make_variogram([], nugget=1.0)
    e = oracle.block_krige_system(d, pure, oracle.block_points([0, 0, 0], _abi.make_block((l, 1, 1), (400, 1, 1))))
    assert e["support"] == 0.0 and abs(e["variance"] - 0.5) < 1e-12


def test_one_point_block_is_point_kriging():
    vg = _abi.make_variogram([("spherical", 0.7, 100.0)], nugget=0.3)
    one = _abi.make_block((10, 10, 10), (1, 1, 1))
    d = samp([(50, 0)], [1])
    p = oracle.krige_system(d, vg, (0, 0, 0))
    b = oracle.block_krige_system(d, vg, oracle.block_points([0, 0, 0], one))
    assert abs(p["variance"] - 2 * oracle.gamma(vg, 50.0)) < 1e-9
    assert abs(b["variance"] - p["variance"]) < 1e-9 and b["support"] == oracle.total_sill(vg)


def test_block_matches_numpy_solve():
    vg = _abi.make_variogram([("gaussian", 1.0, 100.0)], nugget=1e-2)
    rng = np.random.default_rng(5)
    P = rng.uniform(-80, 80, (30, 3))
    d = oracle.Data(P[:, 0], P[:, 1], P[:, 2], rng.uniform(0, 2, 30))
    pts = oracle.block_points([3, -2, 1], _abi.make_block((10, 10, 5), (4, 4, 2)))
    n = 30
    A = np.ones((n + 1, n + 1))
    A[n, n] = 0
    for i in range(n):
        for j in range(n):
            A[i, j] = oracle.cov_points(vg, P[i], P[j])
    b = np.ones(n + 1)
    b[:n] = [np.mean([oracle.block_cov_points(vg, P[i], q) for q in pts]) for i in range(n)]
    x = np.linalg.solve(A, b)
    cbb = np.mean([[oracle.block_cov_points(vg, p, q) for q in pts] for p in pts])
    e = oracle.block_krige_system(d, vg, pts)
    assert np.allclose(e["weights"], x[:n], rtol=1e-7, atol=1e-9)
    assert abs(e["support"] - cbb) < 1e-13
    assert abs(e["variance"] - max(cbb - x[:n] @ b[:n] - x[n], 0)) < 1e-9
