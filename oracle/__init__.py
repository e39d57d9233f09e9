"""The oracle: a CPU restatement of boitatá's kriging hot path.

TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py may load it, and only as the checker or as the
timed CPU baseline. The engine (boitata_amd) never imports it.

oracle.cpp holds the restatement (each function cites the reference file and
line it follows); this module binds liboracle.so with ctypes.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

from boitata_amd import _abi

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


class Samples(C.Structure):
    _fields_ = [
        ("n", C.c_int64),
        ("x", C.c_void_p),
        ("y", C.c_void_p),
        ("z", C.c_void_p),
        ("value", C.c_void_p),
        ("err", C.c_void_p),
        ("hole", C.c_void_p),
    ]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        P, D, I64, I32 = C.c_void_p, C.c_double, C.c_int64, C.c_int
        SP, VP, SR, BP = C.POINTER(Samples), C.POINTER(_abi.Variogram), C.POINTER(_abi.Search), C.POINTER(_abi.Block)
        sig = {
            "or_rotation_matrix": (None, [D, D, D, P]),
            "or_block_frame": (None, [P, P]),
            "or_aniso_matrix": (None, [D, D, D, D, D, D, P]),
            "or_shape": (D, [I32, D, D, D]),
            "or_d_shape": (None, [I32, D, D, P]),
            "or_cov": (D, [VP, D]),
            "or_gamma": (D, [VP, D]),
            "or_total_sill": (D, [VP]),
            "or_cov_points": (D, [VP, P, P]),
            "or_block_cov_points": (D, [VP, P, P]),
            "or_check_search": (I32, [SR]),
            "or_block_points": (None, [P, BP, P]),
            "or_krige_system": (I32, [SP, VP, I32, D, P, I32, P, P, P, P, P]),
            "or_block_krige_system": (I32, [SP, VP, I32, D, P, I64, P, P, P, P, P, P]),
            "or_neighbors": (I32, [SP, SR, VP, I64, P, P, P, P, I32, P, P, P, I32]),
            "or_krige": (I32, [SP, VP, SR, I32, D, I64, P, P, P, P, P, P, P, P, I32]),
            "or_block_krige": (I32, [SP, VP, SR, I32, D, BP, I64, P, P, P, P, P, P, P, P, P, I32]),
            "or_loo": (I32, [SP, VP, SR, I32, D, I64, P, P, P, P, I32]),
            "or_krige_lists": (I32, [SP, VP, I32, D, I64, P, P, P, P, P, I64, P, P, P, P, I32]),
            "or_block_krige_lists": (I32, [SP, VP, I32, D, BP, I64, P, P, P, P, P, I64, I32, P, P, P, P, P, P,
                                           I32]),
            "or_system_matrix": (None, [SP, VP, I32, P, P]),
            "or_sgs": (I32, [SP, VP, SR, I64, P, P, P, P, P, P, P]),
            "or_sgs_plan": (I32, [SP, VP, SR, I64, P, P, P, P, P, P, P, I32]),
            "or_prefix_quality": (I32, [SP, VP, I32, P, BP, I32, P, P, P]),
            "or_calibrated_count": (I32, [I32, D, I64, I32, P, P, P]),
            "or_balanced": (I32, [SP, SR, VP, P, I32, P, P]),
            "or_neighborhood_stats": (None, [SP, P, I32, P, I64, D, P]),
            "or_krige_diag": (I32, [SP, VP, SR, I32, D, P, I64, P, P, P, C.POINTER(_abi.DiagOut), I32]),
            "or_mik_predict": (I32, [SP, C.POINTER(_abi.Mik), P, P, P, SR, P, I64, P, P, P, I32, P, I32, P,
                                     C.POINTER(_abi.MikOut), I32]),
            "or_mik_cross_validate": (I32, [SP, C.POINTER(_abi.Mik), P, P, P, SR, I64, P, P,
                                            C.POINTER(_abi.MikOut), I32]),
            "or_mik_global": (I32, [I64, P, P, I32, P, P, P, P, P, P, P]),
            "or_mik_shape": (D, [I64, P, P, I32, P, P, I32, D, I64, I64, D, D, D, I32]),
            "or_mik_conditional": (I32, [I64, P, P, C.POINTER(_abi.Mik), P, I64, P, I32, P, I32, P,
                                         C.POINTER(_abi.MikOut)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


class Data:
    """Samples held for the oracle (SoA, float64 / int32)."""

    def __init__(self, x, y, z, value=None, err=None, hole=None):
        self.x = np.ascontiguousarray(x, dtype=np.float64)
        self.y = np.ascontiguousarray(y, dtype=np.float64)
        self.z = np.ascontiguousarray(z, dtype=np.float64)
        self.value = np.ascontiguousarray(value if value is not None else np.zeros(len(self.x)), dtype=np.float64)
        self.err = None if err is None else np.ascontiguousarray(err, dtype=np.float64)
        self.hole = None if hole is None else np.ascontiguousarray(hole, dtype=np.int32)
        self.s = Samples(len(self.x), _p(self.x), _p(self.y), _p(self.z), _p(self.value), _p(self.err), _p(self.hole))

    @property
    def n(self):
        return len(self.x)


def rotation_matrix(az, dip, rake):
    out = np.zeros(9)
    lib().or_rotation_matrix(az, dip, rake, _p(out))
    return out.reshape(3, 3)


def block_frame(rot):
    r = np.ascontiguousarray(rot, dtype=np.float64)
    out = np.zeros(9)
    lib().or_block_frame(_p(r), _p(out))
    return out.reshape(3, 3)


def aniso_matrix(az, dip, rake, major, semi, minor):
    out = np.zeros(9)
    lib().or_aniso_matrix(az, dip, rake, major, semi, minor, _p(out))
    return out.reshape(3, 3)


def shape(model, h, rng, param=0.0):
    m = _abi.MODELS[model] if isinstance(model, str) else model
    return lib().or_shape(m, param, h, rng)


def d_shape(model, t, param=0.0):
    m = _abi.MODELS[model] if isinstance(model, str) else model
    out = np.zeros(2)
    lib().or_d_shape(m, param, t, _p(out))
    return out


def cov(vg, h):
    return lib().or_cov(C.byref(vg), h)


def gamma(vg, h):
    return lib().or_gamma(C.byref(vg), h)


def total_sill(vg):
    return lib().or_total_sill(C.byref(vg))


def cov_points(vg, p, q):
    p = np.ascontiguousarray(p, dtype=np.float64)
    q = np.ascontiguousarray(q, dtype=np.float64)
    return lib().or_cov_points(C.byref(vg), _p(p), _p(q))


def block_cov_points(vg, p, q):
    p = np.ascontiguousarray(p, dtype=np.float64)
    q = np.ascontiguousarray(q, dtype=np.float64)
    return lib().or_block_cov_points(C.byref(vg), _p(p), _p(q))


def check_search(s):
    return lib().or_check_search(C.byref(s)) == 0


def block_points(t, blk):
    t = np.ascontiguousarray(t, dtype=np.float64)
    out = np.zeros((blk.nx * blk.ny * blk.nz, 3))
    lib().or_block_points(_p(t), C.byref(blk), _p(out))
    return out


def krige_system(data: Data, vg, target, kind=_abi.ORDINARY, param=0.0, solver=0):
    """One system: solver 0 = krige (guard, kernel, LU), 1 = kernel only, 2 = LU only."""
    t = np.ascontiguousarray(target, dtype=np.float64)
    v, var, lag = C.c_double(), C.c_double(), C.c_double()
    nu = C.c_int32()
    w = np.zeros(max(data.n, 1))
    st = lib().or_krige_system(C.byref(data.s), C.byref(vg), kind, param, _p(t), solver, C.byref(v), C.byref(var),
                               C.byref(lag), _p(w), C.byref(nu))
    return dict(status=st, value=v.value, variance=var.value, lagrange=lag.value, weights=w[: nu.value],
                n_used=nu.value)


def block_krige_system(data: Data, vg, pts, kind=_abi.ORDINARY, param=0.0):
    pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
    v, var, lag, sup = C.c_double(), C.c_double(), C.c_double(), C.c_double()
    nu = C.c_int32()
    w = np.zeros(max(data.n, 1))
    st = lib().or_block_krige_system(C.byref(data.s), C.byref(vg), kind, param, _p(pts), len(pts), C.byref(v),
                                     C.byref(var), C.byref(lag), C.byref(sup), _p(w), C.byref(nu))
    return dict(status=st, value=v.value, variance=var.value, lagrange=lag.value, support=sup.value,
                weights=w[: nu.value], n_used=nu.value)


def neighbors(data: Data, search, vg, targets, left_out=None, tree=True, threads=0):
    t = np.ascontiguousarray(targets, dtype=np.float64).reshape(-1, 3)
    tx, ty, tz = (np.ascontiguousarray(t[:, k]) for k in range(3))
    m = len(t)
    lo = None if left_out is None else np.ascontiguousarray(left_out, dtype=np.int64)
    idx = np.full((m, int(search.max_samples) if search.max_samples < 1 << 20 else max(data.n, 1)), -1, np.int32)
    stride_search = search
    if search.max_samples >= 1 << 20:
        raise ValueError("oracle neighbour lists need max_samples < 2^20")
    count = np.zeros(m, np.int32)
    status = np.zeros(m, np.int32)
    vgp = C.byref(vg) if vg is not None else None
    rc = lib().or_neighbors(C.byref(data.s), C.byref(stride_search), vgp, m, _p(tx), _p(ty), _p(tz), _p(lo),
                            1 if tree else 0, _p(idx), _p(count), _p(status), threads)
    if rc != 0:
        raise ValueError("invalid search")
    return idx, count, status


def krige(data: Data, vg, search, targets, kind=_abi.ORDINARY, param=0.0, threads=0):
    t = np.ascontiguousarray(targets, dtype=np.float64).reshape(-1, 3)
    tx, ty, tz = (np.ascontiguousarray(t[:, k]) for k in range(3))
    m = len(t)
    out = {k: np.zeros(m) for k in ("value", "variance", "lagrange")}
    n_used = np.zeros(m, np.int32)
    status = np.zeros(m, np.int32)
    rc = lib().or_krige(C.byref(data.s), C.byref(vg), C.byref(search), kind, param, m, _p(tx), _p(ty), _p(tz),
                        _p(out["value"]), _p(out["variance"]), _p(out["lagrange"]), _p(n_used), _p(status), threads)
    if rc != 0:
        raise ValueError("invalid search")
    out.update(n_used=n_used, status=status)
    return out


def block_krige(data: Data, vg, search, blk, targets, kind=_abi.ORDINARY, param=0.0, threads=0):
    t = np.ascontiguousarray(targets, dtype=np.float64).reshape(-1, 3)
    tx, ty, tz = (np.ascontiguousarray(t[:, k]) for k in range(3))
    m = len(t)
    out = {k: np.zeros(m) for k in ("value", "variance", "lagrange", "support")}
    n_used = np.zeros(m, np.int32)
    status = np.zeros(m, np.int32)
    rc = lib().or_block_krige(C.byref(data.s), C.byref(vg), C.byref(search), kind, param, C.byref(blk), m, _p(tx),
                              _p(ty), _p(tz), _p(out["value"]), _p(out["variance"]), _p(out["lagrange"]),
                              _p(out["support"]), _p(n_used), _p(status), threads)
    if rc != 0:
        raise ValueError("invalid search or block")
    out.update(n_used=n_used, status=status)
    return out


def loo(data: Data, vg, search, which=None, kind=_abi.ORDINARY, param=0.0, threads=0):
    w = None if which is None else np.ascontiguousarray(which, dtype=np.int64)
    m = data.n if w is None else len(w)
    value = np.zeros(m)
    variance = np.zeros(m)
    status = np.zeros(m, np.int32)
    rc = lib().or_loo(C.byref(data.s), C.byref(vg), C.byref(search), kind, param, m, _p(w), _p(value), _p(variance),
                      _p(status), threads)
    if rc != 0:
        raise ValueError("invalid search")
    return dict(value=value, variance=variance, status=status)


def krige_lists(data: Data, vg, targets, idx, count, kind=_abi.ORDINARY, param=0.0, threads=0):
    t = np.ascontiguousarray(targets, dtype=np.float64).reshape(-1, 3)
    tx, ty, tz = (np.ascontiguousarray(t[:, k]) for k in range(3))
    idx = np.ascontiguousarray(idx, dtype=np.int32)
    count = np.ascontiguousarray(count, dtype=np.int32)
    m = len(t)
    value, variance, lagrange = np.zeros(m), np.zeros(m), np.zeros(m)
    status = np.zeros(m, np.int32)
    lib().or_krige_lists(C.byref(data.s), C.byref(vg), kind, param, m, _p(tx), _p(ty), _p(tz), _p(idx), _p(count),
                         idx.shape[1], _p(value), _p(variance), _p(lagrange), _p(status), threads)
    return dict(value=value, variance=variance, lagrange=lagrange, status=status)


def block_krige_lists(data: Data, vg, blk, targets, idx, count, solver=0, lanes=None, kind=_abi.ORDINARY,
                      param=0.0, threads=0):
    """Block kriging of `targets` from the given neighbour lists: solver 0 is
    the reference's LU (block_krige_points), 1 the engine's FP64 order
    (Cholesky, Schur scalar, butterfly sums over lanes[t] lanes) with the LU
    for what that order leaves to it."""
    t = np.ascontiguousarray(targets, dtype=np.float64).reshape(-1, 3)
    tx, ty, tz = (np.ascontiguousarray(t[:, k]) for k in range(3))
    idx = np.ascontiguousarray(idx, dtype=np.int32)
    count = np.ascontiguousarray(count, dtype=np.int32)
    m = len(t)
    ln = None if lanes is None else np.ascontiguousarray(lanes, dtype=np.int32)
    if solver == 1 and ln is None:
        raise ValueError("the engine's order needs the lane group of each system")
    out = {k: np.zeros(m) for k in ("value", "variance", "lagrange", "support")}
    status = np.zeros(m, np.int32)
    lib().or_block_krige_lists(C.byref(data.s), C.byref(vg), kind, param, C.byref(blk), m, _p(tx), _p(ty), _p(tz),
                               _p(idx), _p(count), idx.shape[1], solver, _p(ln), _p(out["value"]),
                               _p(out["variance"]), _p(out["lagrange"]), _p(out["support"]), _p(status), threads)
    out["status"] = status
    return out


def sgs(data: Data, vg, search, grid, path, z, plan=False, threads=0):
    """Sequential Gaussian simulation of scores at the grid nodes (m, 3),
    node by node along `path` (or_sgs; the data's values are their scores),
    or with plan=True as per-node plans over one grid (or_sgs_plan, parallel
    searches and solves). Returns (score, source) per node."""
    g = np.ascontiguousarray(grid, dtype=np.float64).reshape(-1, 3)
    gx, gy, gz = (np.ascontiguousarray(g[:, k]) for k in range(3))
    m = len(g)
    path = np.ascontiguousarray(path, dtype=np.int64)
    z = np.ascontiguousarray(z, dtype=np.float64)
    assert len(path) == m and len(z) >= m
    score = np.full(m, np.nan)
    source = np.zeros(m, np.int32)
    if plan:
        rc = lib().or_sgs_plan(C.byref(data.s), C.byref(vg), C.byref(search), m, _p(gx), _p(gy), _p(gz), _p(path),
                               _p(z), _p(score), _p(source), threads)
    else:
        rc = lib().or_sgs(C.byref(data.s), C.byref(vg), C.byref(search), m, _p(gx), _p(gy), _p(gz), _p(path),
                          _p(z), _p(score), _p(source))
    if rc != 0:
        raise ValueError("invalid search or no data")
    return score, source


def system_matrix(data: Data, vg, idx):
    """The covariance matrix (cov_points, error variances on the diagonal) of the samples idx."""
    idx = np.ascontiguousarray(idx, dtype=np.int32)
    out = np.zeros((len(idx), len(idx)))
    lib().or_system_matrix(C.byref(data.s), C.byref(vg), len(idx), _p(idx), _p(out))
    return out


def prefix_quality(data: Data, vg, target, idx, kind=_abi.ORDINARY, block=None):
    """(slope, efficiency) of kriging `target` from each leading part of the
    samples idx (in that order); NaN where the reference gives None."""
    idx = np.ascontiguousarray(idx, dtype=np.int64)
    t = np.ascontiguousarray(target, dtype=np.float64)
    n = len(idx)
    slope, eff = np.zeros(n), np.zeros(n)
    lib().or_prefix_quality(C.byref(data.s), C.byref(vg), kind, _p(t), C.byref(block) if block is not None else None,
                            n, _p(idx), _p(slope), _p(eff))
    return slope, eff


def calibrated_count(calibration, target, min_samples, slope, efficiency):
    slope = np.ascontiguousarray(slope, dtype=np.float64)
    efficiency = np.ascontiguousarray(efficiency, dtype=np.float64)
    met = C.c_int32()
    n = lib().or_calibrated_count(int(calibration), float(target), int(min_samples), len(slope), _p(slope),
                                  _p(efficiency), C.byref(met))
    return n, bool(met.value)


def balanced(data: Data, search, vg, target, chosen):
    chosen = np.ascontiguousarray(chosen, dtype=np.int64)
    t = np.ascontiguousarray(target, dtype=np.float64)
    out = np.zeros(len(chosen), np.int64)
    rc = lib().or_balanced(C.byref(data.s), C.byref(search), C.byref(vg) if vg is not None else None, _p(t),
                           len(chosen), _p(chosen), _p(out))
    if rc != 0:
        raise ValueError("invalid search")
    return out


def _variogram_array(variograms):
    vs = list(variograms) if isinstance(variograms, (list, tuple)) else [variograms]
    arr = (_abi.Variogram * len(vs))(*vs)
    return arr, len(vs)


def _mik_out(m, k, nc, nq):
    out = dict(raw=np.zeros((k, m)), cdf=np.zeros((k, m)), correction=np.zeros(m), mean=np.zeros(m),
               variance=np.zeros(m), probability_above=np.zeros((nc, m)), mean_above=np.zeros((nc, m)),
               quantile_values=np.zeros((nq, m)), violations=np.zeros(m, np.int32), n_used=np.zeros(m, np.int32),
               status=np.zeros(m, np.int32))
    o = _abi.MikOut(*[out[name].ctypes.data for name, _ in _abi.MikOut._fields_])
    return out, o


def mik_predict(data: Data, mik, thresholds, variograms, search, targets, weights=None, block=None, cutoffs=(),
                quantiles=(), threads=0):
    """MultipleIndicator::predict for one pass; outputs entry-major (k, m)."""
    t = np.ascontiguousarray(targets, dtype=np.float64).reshape(-1, 3)
    tx, ty, tz = (np.ascontiguousarray(t[:, k]) for k in range(3))
    thr = np.ascontiguousarray(thresholds, dtype=np.float64)
    w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
    cut = np.ascontiguousarray(cutoffs, dtype=np.float64)
    qs = np.ascontiguousarray(quantiles, dtype=np.float64)
    vgs, _ = _variogram_array(variograms)
    out, o = _mik_out(len(t), len(thr), len(cut), len(qs))
    rc = lib().or_mik_predict(C.byref(data.s), C.byref(mik), _p(thr), C.cast(vgs, C.c_void_p), _p(w),
                              C.byref(search), C.byref(block) if block is not None else None, len(t), _p(tx), _p(ty),
                              _p(tz), len(cut), _p(cut), len(qs), _p(qs), C.byref(o), threads)
    if rc != 0:
        raise ValueError(f"invalid parameters ({rc})")
    return out


def mik_cross_validate(data: Data, mik, thresholds, variograms, search, weights=None, which=None, threads=0):
    thr = np.ascontiguousarray(thresholds, dtype=np.float64)
    w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
    wh = None if which is None else np.ascontiguousarray(which, dtype=np.int64)
    m = data.n if wh is None else len(wh)
    vgs, _ = _variogram_array(variograms)
    out, o = _mik_out(m, len(thr), 0, 0)
    pit = np.zeros(m)
    rc = lib().or_mik_cross_validate(C.byref(data.s), C.byref(mik), _p(thr), C.cast(vgs, C.c_void_p), _p(w),
                                     C.byref(search), m, _p(wh), _p(pit), C.byref(o), threads)
    if rc != 0:
        raise ValueError(f"invalid parameters ({rc})")
    out["pit"] = pit
    return out


def mik_global(values, thresholds, weights=None, tails=None):
    v = np.ascontiguousarray(values, dtype=np.float64)
    thr = np.ascontiguousarray(thresholds, dtype=np.float64)
    w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
    tl = None if tails is None else np.ascontiguousarray(tails, dtype=np.float64)
    n, k = len(v), len(thr)
    sorted_, sums = np.zeros(n), np.zeros((3, n + 1))
    classes, bounds, props = np.zeros((k + 1, 2), np.int64), np.zeros(2), np.zeros(k)
    st = lib().or_mik_global(n, _p(v), _p(w), k, _p(thr), _p(tl), _p(sorted_), _p(sums), _p(classes), _p(bounds),
                             _p(props))
    if st != 0:
        raise ValueError(f"Global::new failed ({st})")
    return dict(values=sorted_, sums=sums, classes=classes, lower=bounds[0], upper=bounds[1], proportions=props)


SHAPES = {"data": 0, "power": 1, "hyperbolic": 2}


def mik_shape(values, thresholds, shape, a, b, c, j=None, u=None, w=1.0, cls=(0, 0), weights=None, tails=None):
    """Shape::above(c, j) or, with u, Shape::quantile(u) of one class."""
    v = np.ascontiguousarray(values, dtype=np.float64)
    thr = np.ascontiguousarray(thresholds, dtype=np.float64)
    wt = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
    tl = None if tails is None else np.ascontiguousarray(tails, dtype=np.float64)
    jj, cc = (-1, float(u)) if u is not None else (int(j), float(c))
    return lib().or_mik_shape(len(v), _p(v), _p(wt), len(thr), _p(thr), _p(tl), SHAPES[shape], float(w), int(cls[0]),
                              int(cls[1]), float(a), float(b), cc, jj)


def mik_conditional(values, mik, thresholds, raw, cutoffs=(), quantiles=(), weights=None):
    """MultipleIndicator::conditional of kriged vectors raw (k, m)."""
    v = np.ascontiguousarray(values, dtype=np.float64)
    thr = np.ascontiguousarray(thresholds, dtype=np.float64)
    r = np.ascontiguousarray(raw, dtype=np.float64).reshape(len(thr), -1)
    wt = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
    cut = np.ascontiguousarray(cutoffs, dtype=np.float64)
    qs = np.ascontiguousarray(quantiles, dtype=np.float64)
    m = r.shape[1]
    out, o = _mik_out(m, len(thr), len(cut), len(qs))
    rc = lib().or_mik_conditional(len(v), _p(v), _p(wt), C.byref(mik), _p(thr), m, _p(r), len(cut), _p(cut), len(qs),
                                  _p(qs), C.byref(o))
    if rc != 0:
        raise ValueError(f"invalid parameters ({rc})")
    return out


STATS = ("n_within", "n_holes", "nearest_dist", "mean_dist_knn", "max_dist_knn", "value_mean", "value_var",
         "value_min", "value_max", "value_idw")


def neighborhood_stats(data: Data, target, idx, k, radius=float("inf")):
    """neighborhood_stats of the samples idx at target (Euclidean distance)."""
    idx = np.ascontiguousarray(idx, dtype=np.int64)
    t = np.ascontiguousarray(target, dtype=np.float64)
    out = np.zeros(10)
    lib().or_neighborhood_stats(C.byref(data.s), _p(t), len(idx), _p(idx), int(k), float(radius), _p(out))
    return dict(zip(STATS, out))


DIAG_F = ("value", "variance", "lagrange", "support_variance", "negative_weight_sum", "mean_distance")
DIAG_I = ("n_used", "status", "n_holes", "target_met")


def krige_diag(data: Data, vg, search, targets, kind=_abi.ORDINARY, param=0.0, block=None, threads=0):
    """One pass of point (block None) or block kriging with the per-target diagnostics."""
    t = np.ascontiguousarray(targets, dtype=np.float64).reshape(-1, 3)
    tx, ty, tz = (np.ascontiguousarray(t[:, k]) for k in range(3))
    m = len(t)
    out = {k: np.zeros(m) for k in DIAG_F}
    out.update({k: np.zeros(m, np.int32) for k in DIAG_I})
    o = _abi.DiagOut(*[out[name].ctypes.data for name, _ in _abi.DiagOut._fields_])
    rc = lib().or_krige_diag(C.byref(data.s), C.byref(vg), C.byref(search), kind, param,
                             C.byref(block) if block is not None else None, m, _p(tx), _p(ty), _p(tz), C.byref(o),
                             threads)
    if rc != 0:
        raise ValueError("invalid search")
    return out
