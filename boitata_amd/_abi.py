"""ctypes mirror of include/boitata_gpu.h (structs, enums, signatures).

Pure Python: importing it loads no library.
"""
from __future__ import annotations

import ctypes as C

# Per-target status (crates/estimation/src/error.rs:4-16).
OK = 0
SEARCH_FAILED = 1
SINGULAR = 2
INSUFFICIENT_DATA = 3
INVALID_PARAMETERS = 4
UNSUPPORTED = 5

E_INVALID = -1
E_HIP = -2
E_STATE = -3
E_NOMEM = -4

ORDINARY, SIMPLE, INDICATOR = 0, 1, 2
FP64, FP32_IR = 0, 1

MODELS = {
    "spherical": 0,
    "exponential": 1,
    "gaussian": 2,
    "cubic": 3,
    "pentaspherical": 4,
    "circular": 5,
    "sinehole": 6,
    "matern": 7,
    "power": 8,
}
MAX_STRUCTURES = 4
MAX_CHOL = 64
MAX_LU = 79


class Structure(C.Structure):
    _fields_ = [
        ("model", C.c_int32),
        ("reserved", C.c_int32),
        ("param", C.c_double),
        ("sill", C.c_double),
        ("range", C.c_double),
        ("ratio_semi", C.c_double),
        ("ratio_minor", C.c_double),
    ]


class Variogram(C.Structure):
    _fields_ = [
        ("nugget", C.c_double),
        ("n_structures", C.c_int32),
        ("has_anisotropy", C.c_int32),
        ("structures", Structure * MAX_STRUCTURES),
        ("azimuth", C.c_double),
        ("dip", C.c_double),
        ("rake", C.c_double),
        ("ratio_semi", C.c_double),
        ("ratio_minor", C.c_double),
    ]


class Search(C.Structure):
    _fields_ = [
        ("min_samples", C.c_int64),
        ("max_samples", C.c_int64),
        ("radius", C.c_double),
        ("has_anisotropy", C.c_int32),
        ("octant", C.c_int32),
        ("azimuth", C.c_double),
        ("dip", C.c_double),
        ("rake", C.c_double),
        ("ratio_semi", C.c_double),
        ("ratio_minor", C.c_double),
        ("max_per_hole", C.c_int64),
        ("max_per_octant", C.c_int64),
        ("sectors", C.c_int64),
        ("max_per_sector", C.c_int64),
        ("min_sectors", C.c_int64),
        ("max_empty_sectors", C.c_int64),
        ("min_holes", C.c_int64),
        ("calibration", C.c_int32),
        ("reserved", C.c_int32),
        ("calibration_target", C.c_double),
    ]


class Block(C.Structure):
    _fields_ = [
        ("size", C.c_double * 3),
        ("nx", C.c_int32),
        ("ny", C.c_int32),
        ("nz", C.c_int32),
        ("reserved", C.c_int32),
        ("rotation", C.c_double * 3),
    ]


def make_variogram(structures, nugget=0.0, rotation=None, ratios=(1.0, 1.0)) -> Variogram:
    """structures: [(model, sill, range[, param[, (semi, minor)]])]."""
    v = Variogram()
    v.nugget = float(nugget)
    v.n_structures = len(structures)
    if len(structures) > MAX_STRUCTURES:
        raise ValueError("at most 4 structures")
    for k, s in enumerate(structures):
        model, sill, rng = s[0], s[1], s[2]
        param = s[3] if len(s) > 3 and s[3] is not None else 0.0
        own = s[4] if len(s) > 4 else None
        st = v.structures[k]
        st.model = MODELS[model] if isinstance(model, str) else int(model)
        st.param = float(param)
        st.sill = float(sill)
        st.range = float(rng)
        st.ratio_semi = float(own[0]) if own is not None else 0.0
        st.ratio_minor = float(own[1]) if own is not None else 0.0
    rot = tuple(rotation) if rotation is not None else (0.0, 0.0, 0.0)
    rat = tuple(ratios)
    # crates/py/src/variogram.rs:164-177: the identity is no anisotropy.
    if rot != (0.0, 0.0, 0.0) or rat != (1.0, 1.0):
        v.has_anisotropy = 1
        v.azimuth, v.dip, v.rake = (float(a) for a in rot)
        v.ratio_semi, v.ratio_minor = (float(r) for r in rat)
    return v


CALIBRATE_SLOPE = 1
CALIBRATE_EFFICIENCY = 2


def make_search(min_samples=4, max_samples=16, radius=float("inf"), rotation=None, ratios=None, octant=False,
                max_per_hole=None, max_per_octant=None, sectors=None, max_per_sector=None, min_sectors=None,
                max_empty_sectors=None, min_holes=None, target_slope=None, target_efficiency=None) -> Search:
    s = Search()
    s.min_samples = int(min_samples)
    s.max_samples = int(max_samples)
    s.radius = float(radius)
    if rotation is not None or ratios is not None:
        rot = tuple(rotation) if rotation is not None else (0.0, 0.0, 0.0)
        rat = tuple(ratios) if ratios is not None else (1.0, 1.0)
        if rot != (0.0, 0.0, 0.0) or rat != (1.0, 1.0):
            s.has_anisotropy = 1
            s.azimuth, s.dip, s.rake = (float(a) for a in rot)
            s.ratio_semi, s.ratio_minor = (float(r) for r in rat)
    s.octant = 1 if octant else 0
    s.max_per_hole = -1 if max_per_hole is None else int(max_per_hole)
    s.max_per_octant = -1 if max_per_octant is None else int(max_per_octant)
    s.sectors = 0 if sectors is None else int(sectors)
    s.max_per_sector = 0 if max_per_sector is None else int(max_per_sector)
    s.min_sectors = -1 if min_sectors is None else int(min_sectors)
    s.max_empty_sectors = -1 if max_empty_sectors is None else int(max_empty_sectors)
    s.min_holes = -1 if min_holes is None else int(min_holes)
    if target_slope is not None:
        s.calibration, s.calibration_target = CALIBRATE_SLOPE, float(target_slope)
    elif target_efficiency is not None:
        s.calibration, s.calibration_target = CALIBRATE_EFFICIENCY, float(target_efficiency)
    return s


def make_block(size, discretization=(4, 4, 1), rotation=(0.0, 0.0, 0.0)) -> Block:
    b = Block()
    for d in range(3):
        b.size[d] = float(size[d])
        b.rotation[d] = float(rotation[d])
    b.nx, b.ny, b.nz = (int(k) for k in discretization)
    return b


class DiagOut(C.Structure):
    """bk_diag_out: output pointers of bk_krige_diagnostics; NULL ones are skipped."""

    _fields_ = [(name, C.c_void_p) for name in ("value", "variance", "lagrange", "support_variance", "n_used",
                                                 "status", "negative_weight_sum", "n_holes", "mean_distance",
                                                 "target_met")]


MIK_GLOBAL, MIK_LINEAR = 0, 1
TAIL_NONE, TAIL_POWER, TAIL_HYPERBOLIC = 0, 1, 2


class Mik(C.Structure):
    """bk_mik: multiple indicator kriging parameters (indicator.rs:53-67)."""

    _fields_ = [
        ("n_thresholds", C.c_int32),
        ("n_variograms", C.c_int32),
        ("simple", C.c_int32),
        ("interpolation", C.c_int32),
        ("upper_tail", C.c_int32),
        ("has_tails", C.c_int32),
        ("omega", C.c_double),
        ("tails", C.c_double * 2),
    ]


class MikOut(C.Structure):
    """bk_mik_out: output pointers (entry-major [k * m + t]); NULL ones are skipped."""

    _fields_ = [(name, C.c_void_p) for name in ("raw", "cdf", "correction", "mean", "variance", "probability_above",
                                                 "mean_above", "quantile_values", "violations", "n_used", "status")]


def make_mik(thresholds, n_variograms=1, simple=False, interpolation=MIK_GLOBAL, upper_tail=TAIL_NONE, omega=1.0,
             tails=None) -> Mik:
    p = Mik()
    p.n_thresholds = len(thresholds)
    p.n_variograms = int(n_variograms)
    p.simple = 1 if simple else 0
    p.interpolation = int(interpolation)
    p.upper_tail = int(upper_tail)
    p.omega = float(omega)
    p.has_tails = 0 if tails is None else 1
    if tails is not None:
        p.tails[0], p.tails[1] = float(tails[0]), float(tails[1])
    return p


P = C.c_void_p
I64 = C.c_int64
I32 = C.c_int32
D = C.c_double

# (name, restype, argtypes) of every exported symbol.
SIGNATURES = [
    ("bk_create", C.c_int, [C.POINTER(C.c_void_p), C.c_int]),
    ("bk_destroy", None, [P]),
    ("bk_last_error", C.c_char_p, [P]),
    ("bk_version", C.c_char_p, []),
    ("bk_check_search", C.c_int, [P, C.POINTER(Search)]),
    ("bk_check_variogram", C.c_int, [P, C.POINTER(Variogram)]),
    ("bk_set_samples", C.c_int, [P, I64, P, P, P, P, P, P]),
    ("bk_neighbors", C.c_int, [P, C.POINTER(Search), C.POINTER(Variogram), I64, P, P, P, P, P, P, P]),
    ("bk_krige", C.c_int, [P, C.POINTER(Variogram), C.POINTER(Search), I32, D, I32, I64, P, P, P, P, P, P, P, P]),
    ("bk_krige_noalloc", C.c_int,
     [P, C.POINTER(Variogram), C.POINTER(Search), I32, D, I32, I64, P, P, P, P, P, P, P, P]),
    ("bk_block_krige", C.c_int,
     [P, C.POINTER(Variogram), C.POINTER(Search), I32, D, C.POINTER(Block), I32, I64, P, P, P, P, P, P, P, P, P]),
    ("bk_loo", C.c_int, [P, C.POINTER(Variogram), C.POINTER(Search), I32, D, I32, I64, P, P, P, P]),
    ("bk_block_loo", C.c_int,
     [P, C.POINTER(Variogram), C.POINTER(Search), I32, D, C.POINTER(Block), I32, I64, P, P, P, P]),
    ("bk_krige_weights", C.c_int,
     [P, C.POINTER(Variogram), C.POINTER(Search), I32, D, P, I32, I64, P, P, P, P, P, P, P]),
    ("bk_sgs", C.c_int, [P, C.POINTER(Variogram), C.POINTER(Search), I32, I64, P, P, P, P, P, P, P]),
    ("bk_loo_prepare", C.c_int, [P, C.POINTER(Search), C.POINTER(Variogram)]),
    ("bk_loo_loss_grad", C.c_int, [P, C.POINTER(Variogram), I32, P, P, P]),
    ("bk_krige_diagnostics", C.c_int,
     [P, C.POINTER(Variogram), C.POINTER(Search), I32, D, P, I32, I64, P, P, P, C.POINTER(DiagOut)]),
    ("bk_mik_predict", C.c_int,
     [P, C.POINTER(Mik), P, P, P, C.POINTER(Search), P, I32, I64, P, P, P, I32, P, I32, P, C.POINTER(MikOut)]),
    ("bk_mik_cross_validate", C.c_int,
     [P, C.POINTER(Mik), P, P, P, C.POINTER(Search), I32, I64, P, P, C.POINTER(MikOut)]),
]

# Declared in include/boitata_gpu.h beside the reference-facing entry points.
EXTRA_SIGNATURES = [
    ("bk_set_tuning", C.c_int, [P, P, I32, D, D]),
    ("bk_profile", C.c_int, [P, I32]),
    ("bk_profile_read", C.c_int, [P, P, P]),
    ("bk_set_search_tuning", C.c_int, [P, I32]),
    ("bk_get_stat", C.c_int, [P, I32, P]),
    ("bk_set_solve_tuning", C.c_int, [P, I32]),
    ("bk_autotune", C.c_int, [P, P, P, I32, D, I32, I64, P, P, P, P]),
    ("bk_set_graphs", C.c_int, [P, I32]),
    ("bk_set_refine_limit", C.c_int, [P, I32]),
]

AUTOTUNE_REPORT = ("cell_factor", "search_mode", "waves", "edge0", "edge1", "edge2", "search_ms_before",
                   "solve_ms_before", "search_ms_after", "solve_ms_after", "refine_limit")

STATIC_MAX_TARGETS = 65536
STATIC_MAX_SAMPLES = 262144
STATIC_MAX_NEIGHBOURS = 64
STAT_SEARCH_SPILL = 0
STAT_GRAPH_LAUNCHES = 1
STAT_REFINED = 2
STAT_REFINE_SWEEPS = 3
STAT_REFINE_REDO = 4
STAT_SEARCH_RETRY = 5
STAT_SGS_PASSES = 6

STAGES = ("grid", "search", "bucket", "solve", "deferred", "total")
