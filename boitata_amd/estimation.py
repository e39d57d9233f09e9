"""Drop-in estimators on the MI355X engine.

The classes keep the constructor, ``fit``, ``predict`` and ``cross_validate``
signatures of the reference's Python API (python/boitata/estimation.py:251-485,
through ``_Estimator``, crates/py/src/estimation.rs:1213-1866) for the methods
the engine covers: ``OrdinaryKriging``, ``SimpleKriging``,
``IndicatorKriging``, ``BlockKriging`` and ``MultipleIndicatorKriging``
(crates/py/src/indicator.rs:116-420); and of ``Variogram``
(crates/py/src/variogram.rs:329-362), ``Structure`` (:211) and ``Search``
(crates/py/src/estimation.rs:415-531).

Behaviour kept from the reference:

* ``fit`` keeps the first of samples sharing a location bit for bit (after
  ``+ 0.0``), with a ``UserWarning`` naming their holes or rows
  (crates/py/src/args.rs:305-363, crates/eda/src/lib.rs:952 at tolerance 0);
  hole labels become integer codes in order of first appearance, an unlabelled
  sample gets a code of its own (args.rs:278-292).
* A sequence of searches is a list of passes: targets a pass leaves
  unestimated (any per-target failure) go to the next (batch.rs:132-147).
* Per-target failures are NaN; parameters out of range raise InvalidInput.
* ``(n, 2)`` coordinates lie on z = 0; coordinates must be finite
  (crates/py/src/containers.rs:15-21).

Not covered (raise NotImplementedError): domains, local anisotropy,
units, soft / high-grade data, and the
other estimation methods.

Extension: ``precision="fp32_ir"`` solves in FP32 with FP64 refinement; the
default ``"fp64"`` is the reference's arithmetic.
"""
from __future__ import annotations

import math
import warnings
from collections.abc import Sequence
from dataclasses import dataclass

import numpy as np

from . import _abi
from .errors import FileError, InvalidInput

__all__ = [
    "BlockKriging",
    "CrossValidation",
    "FileError",
    "IndicatorCrossValidation",
    "IndicatorKriging",
    "IndicatorSummary",
    "InvalidInput",
    "MultipleIndicatorKriging",
    "OrdinaryKriging",
    "Search",
    "SimpleKriging",
    "Structure",
    "Variogram",
]

_PRECISIONS = {"fp64": _abi.FP64, "fp32_ir": _abi.FP32_IR}


def _lib():
    from ._lib import load

    return load()


def _checked(fn, obj, what):
    lib = _lib()
    import ctypes as C

    if fn(lib)(None, C.byref(obj)) != 0:
        raise InvalidInput(f"{what}: {lib.bk_last_error(None).decode()}")


def _number(v, what):
    if isinstance(v, str):
        raise NotImplementedError(f"{what} with a unit is not supported by this engine")
    return float(v)


class Structure:
    """One nested structure: `sill` is the partial sill, `range` the range
    along the major axis; `order` for matern, `exponent` for power
    (crates/py/src/variogram.rs:211). `ratios` (extension) gives the structure
    its own (semi/major, minor/major) ratios under the variogram's rotation."""

    def __init__(self, model: str, sill: float, range, *, order=None, exponent=None, ratios=None):
        if model not in _abi.MODELS:
            raise InvalidInput(f"unknown model {model!r}; one of {', '.join(_abi.MODELS)}")
        if order is not None and model != "matern":
            raise InvalidInput("order is for the matern model")
        if exponent is not None and model != "power":
            raise InvalidInput("exponent is for the power model")
        if model == "matern" and order is None:
            order = 0.5
        if model == "power" and exponent is None:
            exponent = 1.0
        self.model = model
        self.sill = float(sill)
        self.range = _number(range, "range")
        self.param = order if model == "matern" else exponent
        self.ratios = None if ratios is None else (float(ratios[0]), float(ratios[1]))

    def _tuple(self):
        return (self.model, self.sill, self.range, self.param, self.ratios)

    def __repr__(self):
        return f"Structure({self.model!r}, {self.sill}, {self.range})"


class Variogram:
    """Nugget plus nested structures sharing one anisotropy: `rotation` is
    (azimuth, dip, rake) in degrees, `ratios` (semi/major, minor/major)
    (crates/py/src/variogram.rs:329-362). Structures are Structure objects or
    (model, sill, range) tuples."""

    def __init__(self, structures, *, nugget=0.0, rotation=(0.0, 0.0, 0.0), ratios=(1.0, 1.0), length_unit=None):
        if length_unit is not None:
            raise NotImplementedError("length units are not supported by this engine")
        nugget = float(nugget)
        if not (nugget >= 0.0 and math.isfinite(nugget)):
            raise InvalidInput("nugget must be finite and >= 0")
        parsed = []
        for s in structures:
            if isinstance(s, Structure):
                parsed.append(s)
                continue
            try:
                name, sill, rng = s
            except (TypeError, ValueError):
                raise InvalidInput("structures are Structure or (model, sill, range) tuples") from None
            parsed.append(Structure(name, sill, rng))
        if len(parsed) > _abi.MAX_STRUCTURES:
            raise InvalidInput(f"at most {_abi.MAX_STRUCTURES} structures")
        self.structures = parsed
        self.nugget = nugget
        self.rotation = tuple(float(a) for a in rotation)
        self.ratios = tuple(float(r) for r in ratios)
        self._c = _abi.make_variogram([s._tuple() for s in parsed], nugget=nugget, rotation=self.rotation,
                                      ratios=self.ratios)
        _checked(lambda lib: lib.bk_check_variogram, self._c, "variogram")

    @property
    def sill(self) -> float:
        return self.nugget + sum(s.sill for s in self.structures)

    @property
    def has_sill(self) -> bool:
        return all(s.model != "power" for s in self.structures)

    def __repr__(self):
        return f"Variogram({self.structures!r}, nugget={self.nugget})"

    def to_json(self) -> str:
        """JSON of the model (crates/py/src/variogram.rs:569-579)."""
        from . import persist

        return persist.envelope("Variogram", persist.variogram_fields(self))

    @staticmethod
    def from_json(text: str) -> "Variogram":
        """Reads `to_json` output; raises InvalidInput on another class's JSON
        or a newer format."""
        from . import persist

        return persist.variogram_from(persist.open_envelope("Variogram", text))

    def __reduce__(self):
        return Variogram.from_json, (self.to_json(),)


class Search:
    """Search neighbourhood (crates/py/src/estimation.rs:415-531): `radius`, or
    `ranges` (major, semi, minor) giving radius = major and ratios semi/major,
    minor/major; `max_per_sector` caps octants with `octant=True` or plane
    sectors with `sectors`. `target_slope` or `target_efficiency` calibrates
    the search: each target is kriged from the fewest leading samples, from
    min_samples, that reach it (kriging_algebra.rs:338-441)."""

    def __init__(self, radius=None, *, max_samples=16, min_samples=1, octant=False, sectors=None, max_per_sector=None,
                 max_per_hole=None, rotation=None, ratios=None, ranges=None, min_sectors=None, max_empty_sectors=None,
                 min_holes=None, high_grade=None, soft=None, target_slope=None, target_efficiency=None,
                 length_unit=None):
        for name, v in (("high_grade", high_grade), ("soft", soft), ("length_unit", length_unit)):
            if v is not None:
                raise NotImplementedError(f"{name} is not supported by this engine")
        # py/estimation.rs:458-472
        if target_slope is not None and target_efficiency is not None:
            raise InvalidInput("give at most one of target_slope and target_efficiency")
        if target_slope is not None and not (math.isfinite(float(target_slope)) and float(target_slope) > 0.0):
            raise InvalidInput("target_slope must be finite and > 0")
        if target_efficiency is not None and not (math.isfinite(float(target_efficiency))
                                                  and float(target_efficiency) <= 1.0):
            raise InvalidInput("target_efficiency must be finite and <= 1")
        if radius is not None and ranges is not None:
            raise InvalidInput("give radius or ranges, not both")
        if radius is None and ranges is None:
            raise InvalidInput("give radius or ranges")
        if radius is not None:
            r = _number(radius, "radius")
        else:
            if ratios is not None:
                raise InvalidInput("give ratios with radius, not with ranges")
            major, semi, minor = (float(v) for v in ranges)
            if not (major > 0.0 and semi > 0.0 and minor > 0.0):
                raise InvalidInput("ranges must be > 0")
            r, ratios = major, (semi / major, minor / major)
        for name, v in (("max_samples", max_samples), ("min_samples", min_samples), ("sectors", sectors),
                        ("max_per_sector", max_per_sector), ("max_per_hole", max_per_hole),
                        ("min_sectors", min_sectors), ("max_empty_sectors", max_empty_sectors),
                        ("min_holes", min_holes)):
            if v is not None and (not isinstance(v, (int, np.integer)) or isinstance(v, bool) or v < 0):
                raise InvalidInput(f"{name} must be a non-negative integer")
        if max_per_sector is not None and sectors is None and not octant:
            raise InvalidInput("max_per_sector needs octant or sectors")
        max_per_octant = max_per_sector if (octant and sectors is None) else None
        per_sector = None
        if sectors is not None:
            per_sector = max_per_sector if max_per_sector is not None else -(-int(max_samples) // max(int(sectors), 1))
        self.radius, self.max_samples, self.min_samples = r, int(max_samples), int(min_samples)
        self.octant = bool(octant)
        self.rotation, self.ratios = rotation, ratios
        self._c = _abi.make_search(min_samples=min_samples, max_samples=max_samples, radius=r, rotation=rotation,
                                   ratios=ratios, octant=octant, max_per_hole=max_per_hole,
                                   max_per_octant=max_per_octant, sectors=sectors, max_per_sector=per_sector,
                                   min_sectors=min_sectors, max_empty_sectors=max_empty_sectors,
                                   min_holes=min_holes, target_slope=target_slope,
                                   target_efficiency=target_efficiency)
        if rotation is not None or ratios is not None:
            rot = tuple(rotation) if rotation is not None else (0.0, 0.0, 0.0)
            rat = tuple(ratios) if ratios is not None else (1.0, 1.0)
            if any(not math.isfinite(float(a)) for a in rot) or any(not (float(q) > 0.0) for q in rat):
                raise InvalidInput("anisotropy needs finite angles and ratios > 0")
        _checked(lambda lib: lib.bk_check_search, self._c, "search")

    def __repr__(self):
        return f"Search(radius={self.radius}, max_samples={self.max_samples}, min_samples={self.min_samples})"

    def to_json(self) -> str:
        """JSON of the parameters (crates/py/src/estimation.rs:398-412)."""
        from . import persist

        return persist.envelope("Search", persist.search_fields(self))

    @staticmethod
    def from_json(text: str) -> "Search":
        """Reads `to_json` output; raises InvalidInput on another class's JSON
        or a newer format."""
        from . import persist

        return persist.search_from(persist.open_envelope("Search", text))

    def __reduce__(self):
        return Search.from_json, (self.to_json(),)


@dataclass(frozen=True)
class CrossValidation:
    """Cross-validation results; NaN where a sample had too few neighbours
    (python/boitata/estimation.py:113-149)."""

    actual: np.ndarray
    estimate: np.ndarray
    variance: np.ndarray

    @property
    def error(self) -> np.ndarray:
        return self.estimate - self.actual

    @property
    def mean_error(self) -> float:
        return float(np.nanmean(self.error))

    @property
    def rmse(self) -> float:
        return float(np.sqrt(np.nanmean(self.error ** 2)))

    @property
    def correlation(self) -> float:
        ok = ~np.isnan(self.estimate)
        return float(np.corrcoef(self.actual[ok], self.estimate[ok])[0, 1])

    @property
    def slope(self) -> float:
        ok = ~np.isnan(self.estimate)
        return float(np.polyfit(self.estimate[ok], self.actual[ok], 1)[0])

    @property
    def standardized_squared_error(self) -> float:
        return float(np.nanmean(self.error ** 2 / self.variance))


# --------------------------------------------------------------------------
# Host-side input handling (no device needed).

def coords_of(coords) -> np.ndarray:
    """(n, 3) float64 rows of an (n, 2) or (n, 3) array; (n, 2) lies on z = 0."""
    if hasattr(coords, "to_length_unit") or hasattr(coords, "centroids"):
        raise NotImplementedError("PointSet / BlockModel containers are not supported by this engine; pass arrays")
    try:
        a = np.asarray(coords, dtype=np.float64)
    except (TypeError, ValueError):
        raise InvalidInput("coordinates must be a 2-D array of shape (n, 2) or (n, 3)") from None
    if a.ndim != 2 or a.shape[1] not in (2, 3):
        raise InvalidInput("coordinates must be a 2-D array of shape (n, 2) or (n, 3)")
    if np.isnan(a).any():
        raise InvalidInput("coordinates must be finite")
    if a.shape[1] == 2:
        a = np.concatenate([a, np.zeros((len(a), 1))], axis=1)
    return np.ascontiguousarray(a)


def hole_codes(holes, n):
    """Labels (text) and integer codes of drill holes: codes in order of first
    appearance, a sample without a label gets a fresh code of its own."""
    if holes is None:
        return None, None
    if isinstance(holes, str):
        raise NotImplementedError("hole columns by name need a container; pass an array")
    labels = list(holes)
    if len(labels) != n:
        raise InvalidInput(f"holes: expected {n} values, got {len(labels)}")
    text = [None if (h is None or (isinstance(h, float) and math.isnan(h))) else str(h) for h in labels]
    codes, ids, nxt = {}, np.empty(n, dtype=np.int32), 0
    for i, t in enumerate(text):
        if t is None:
            ids[i] = nxt
            nxt += 1
        else:
            if t not in codes:
                codes[t] = nxt
                nxt += 1
            ids[i] = codes[t]
    return text, ids


def distinct_rows(xyz: np.ndarray, hole_text=None):
    """Rows to keep so that no two samples share a location (the first of each
    group, compared bit for bit after + 0.0), and a warning text naming the
    dropped groups (None when there is none)."""
    keys = (xyz + 0.0).view(np.uint64).reshape(len(xyz), 3)
    groups = {}
    for i, k in enumerate(map(tuple, keys)):
        groups.setdefault(k, []).append(i)
    shared = sorted((g for g in groups.values() if len(g) > 1), key=lambda g: g[0])
    if not shared:
        return np.arange(len(xyz)), None
    dropped = np.zeros(len(xyz), dtype=bool)
    lines = []
    for g in shared:
        dropped[g[1:]] = True
        tagged = None if hole_text is None else [hole_text[i] for i in g]
        if tagged is not None and all(t is not None for t in tagged):
            who = "holes " + ", ".join(tagged)
        else:
            who = "rows " + ", ".join(str(i) for i in g)
        x, y, z = xyz[g[0]]
        lines.append(f"{who} at ({x}, {y}, {z})")
    more = max(len(lines) - 20, 0)
    text = "samples sharing a location keep the first one: " + "; ".join(lines[:20])
    if more:
        text += f"; and {more} more"
    return np.nonzero(~dropped)[0], text


def run_passes(n, searches, one_pass):
    """by_pass (crates/estimation/src/batch.rs:132-147): each search in turn
    on the rows earlier passes left unestimated. one_pass(search, rows) returns
    (value, variance, ok) for those rows. Returns value, variance, the pass of
    each row (-1 where none) as arrays."""
    value = np.full(n, np.nan)
    variance = np.full(n, np.nan)
    which = np.full(n, -1, dtype=np.int64)
    remaining = np.arange(n)
    for p, s in enumerate(searches):
        if len(remaining) == 0:
            break
        v, q, ok = one_pass(s, remaining)
        ok = np.asarray(ok, dtype=bool)
        rows = remaining[ok]
        value[rows] = np.asarray(v)[ok]
        variance[rows] = np.asarray(q)[ok]
        which[rows] = p
        remaining = remaining[~ok]
    return value, variance, which


class DiagnosticsTable:
    """Per-target search diagnostics (crates/py/src/indicator.rs:25-57): named
    float columns, NaN where a target is unestimated."""

    def __init__(self, columns):
        self._columns = {k: np.asarray(v, dtype=np.float64) for k, v in columns}

    @property
    def column_names(self):
        return list(self._columns)

    def __getitem__(self, name):
        return self._columns[name]

    def __len__(self):
        return len(next(iter(self._columns.values()))) if self._columns else 0


def _searches(search):
    if isinstance(search, Search):
        return [search]
    if isinstance(search, Sequence) and len(search) > 0 and all(isinstance(s, Search) for s in search):
        return list(search)
    raise InvalidInput("search must be a Search or a non-empty sequence of Search")


class _Base:
    _kind = _abi.ORDINARY

    def __init__(self, variogram: Variogram, search, *, precision: str = "fp64", device: int = 0):
        if not isinstance(variogram, Variogram):
            raise InvalidInput("kriging needs a Variogram")
        if precision not in _PRECISIONS:
            raise InvalidInput(f"precision is one of {', '.join(_PRECISIONS)}")
        self._variogram = variogram
        self._passes = _searches(search)
        self._precision = _PRECISIONS[precision]
        self._device = device
        self._engine = None
        self._values = None
        self._samples = None  # the fitted samples on the host: x, y, z, value, hole, error_variance
        self._uploaded = False
        self._param = 0.0

    @property
    def variogram(self) -> Variogram:
        return self._variogram

    @property
    def values(self) -> np.ndarray:
        """Values of the fitted samples, after dropping shared locations."""
        if self._values is None:
            raise InvalidInput("call fit first")
        return self._values

    def _device_engine(self):
        if self._engine is None:
            from .engine import Engine

            self._engine = Engine(self._device)
        return self._engine

    def _device_holes(self):
        """Hole codes for the device: a sample without a code (null) counts as a
        hole of its own (crates/estimation/src/neighborhood.rs:85-92); None
        without holes."""
        hole = self._samples["hole"]
        if np.isnan(hole).all():
            return None
        ids = np.empty(len(hole), dtype=np.int32)
        tagged = ~np.isnan(hole)
        ids[tagged] = hole[tagged].astype(np.int32)
        top = int(hole[tagged].max()) + 1 if tagged.any() else 0
        ids[~tagged] = top + np.arange(int((~tagged).sum()), dtype=np.int32)
        return ids

    def _upload(self, eng, rows=None):
        """Sets the fitted samples (all, or those at `rows`) on an engine."""
        s = self._samples
        ids = self._device_holes()
        pick = (lambda a: a) if rows is None else (lambda a: np.ascontiguousarray(a[rows]))
        err = s["error_variance"]
        eng.set_samples(pick(s["x"]), pick(s["y"]), pick(s["z"]), pick(s["value"]),
                        pick(err) if (err != 0.0).any() else None, None if ids is None else pick(ids))

    def _ready(self):
        """The device engine holding the fitted samples (uploaded on first use
        after a restore from a file or a pickle)."""
        eng = self._device_engine()
        if not self._uploaded:
            self._upload(eng)
            self._uploaded = True
        return eng

    def fit(self, coords, values, *, holes=None, error_variance=None, domains=None, domain_column=None):
        """Stores the samples; samples sharing a location keep the first one."""
        if domains is not None or domain_column is not None:
            raise NotImplementedError("domains are not supported by this engine")
        if isinstance(values, str):
            raise NotImplementedError("value columns by name need a container; pass an array")
        xyz = coords_of(coords)
        n = len(xyz)
        v = np.asarray(values, dtype=np.float64).reshape(-1)
        if len(v) != n:
            raise InvalidInput(f"values: expected {n} values, got {len(v)}")
        err = None
        if error_variance is not None:
            err = np.asarray(error_variance, dtype=np.float64).reshape(-1)
            if len(err) != n:
                raise InvalidInput(f"error_variance: expected {n} values, got {len(err)}")
            if not (np.isfinite(err).all() and (err >= 0).all()):
                raise InvalidInput("error_variance must be finite and >= 0")
        text, ids = hole_codes(holes, n)
        keep, note = distinct_rows(xyz, text)
        if note:
            warnings.warn(note, UserWarning, stacklevel=2)
        xyz, v = xyz[keep], v[keep]
        err = None if err is None else err[keep]
        ids = None if ids is None else ids[keep]
        # Host copy of the fitted samples (crates/py/src/estimation.rs:1110-1121):
        # hole codes (an unlabelled sample has a code of its own), null without
        # holes; error variance 0 by default.
        hole = np.full(len(v), np.nan) if ids is None else ids.astype(np.float64)
        self._samples = dict(x=xyz[:, 0].copy(), y=xyz[:, 1].copy(), z=xyz[:, 2].copy(), value=v, hole=hole,
                             error_variance=np.zeros(len(v)) if err is None else err)
        self._values = v
        self._uploaded = False
        self._ready()
        return self

    def _one_pass(self, s, T):
        out = self._engine.krige(self._variogram._c, s._c, T[:, 0].copy(), T[:, 1].copy(), T[:, 2].copy(),
                                 self._kind, self._param, self._precision)
        return out["value"], out["variance"], out["status"] == _abi.OK

    def predict(self, targets, *, return_variance=False, anisotropy=None, diagnostics=False, domains=None,
                domain_column=None):
        """Estimates at targets; NaN where the search found too few samples.
        With `diagnostics` the result is the per-target table of
        crates/py/src/estimation.rs:1046-1100 instead."""
        if anisotropy is not None:
            raise NotImplementedError("local anisotropy is not supported by this engine")
        if domains is not None or domain_column is not None:
            raise NotImplementedError("domains are not supported by this engine")
        if self._values is None:
            raise InvalidInput("call fit first")
        T = coords_of(targets)
        self._ready()
        if diagnostics:
            return self._diagnostics(T)
        value, variance, _ = run_passes(len(T), self._passes, lambda s, rows: self._one_pass(s, T[rows]))
        return (value, variance) if return_variance else value

    _block = None

    def _diagnostics(self, T):
        """py/estimation.rs:1046-1100: the estimate's fields, derived ratios and
        the statistics of the samples used, per target; NaN where unestimated.
        Domains are not supported, so n_other_domain is 0 where estimated."""
        m = len(T)
        cols = {k: np.full(m, np.nan) for k in ("value", "variance", "lagrange", "support_variance", "n_samples",
                                                 "n_holes", "mean_distance", "negative_weight_sum", "target_met",
                                                 "max_samples_reached")}
        which = np.full(m, -1, dtype=np.int64)
        remaining = np.arange(m)
        for p, s in enumerate(self._passes):
            if len(remaining) == 0:
                break
            R = T[remaining]
            out = self._engine.krige_diagnostics(self._variogram._c, s._c, R[:, 0].copy(), R[:, 1].copy(),
                                                 R[:, 2].copy(), self._kind, self._param, self._block,
                                                 self._precision)
            ok = out["status"] == _abi.OK
            rows = remaining[ok]
            for k in ("value", "variance", "lagrange", "support_variance", "negative_weight_sum", "mean_distance"):
                cols[k][rows] = out[k][ok]
            cols["n_samples"][rows] = out["n_used"][ok]
            cols["n_holes"][rows] = out["n_holes"][ok]
            met = out["target_met"][ok].astype(np.float64)
            cols["target_met"][rows] = np.where(met < 0, np.nan, met)
            cols["max_samples_reached"][rows] = (out["n_used"][ok] >= s.max_samples).astype(np.float64)
            which[rows] = p
            remaining = remaining[~ok]
        sv, var, lag = cols["support_variance"], cols["variance"], cols["lagrange"]
        cov = sv - var - lag  # krige.rs:58-61, Cov(Z, Z*)
        est_var = cov - lag   # krige.rs:63-66, Var(Z*)
        ok = which >= 0
        return DiagnosticsTable([
            ("value", cols["value"]),
            ("variance", var),
            ("efficiency", (sv - var) / sv),
            ("slope", cov / est_var),
            ("n_samples", cols["n_samples"]),
            ("pass", np.where(ok, which + 1.0, np.nan)),
            ("n_holes", cols["n_holes"]),
            ("n_other_domain", np.where(ok, 0.0, np.nan)),
            ("mean_distance", cols["mean_distance"]),
            ("negative_weight_sum", cols["negative_weight_sum"]),
            ("lagrange", lag),
            ("support_variance", sv),
            ("estimate_variance", est_var),
            ("max_samples_reached", cols["max_samples_reached"]),
            ("target_met", cols["target_met"]),
        ])

    def _folds(self, k):
        """k_fold_at's folds (crates/estimation/src/batch.rs:217-226): the j-th
        of the sorted distinct hole codes goes to fold j % k, a sample without
        a hole to fold (its index) % k."""
        hole = self._samples["hole"]
        fold = np.arange(len(hole), dtype=np.int64) % k
        tagged = ~np.isnan(hole)
        if tagged.any():
            codes = np.unique(hole[tagged])
            fold[tagged] = np.searchsorted(codes, hole[tagged]) % k
        return fold

    def cross_validate(self, *, folds=None) -> CrossValidation:
        """Estimates and variances at the fitted samples through the same
        passes: leave-one-out, each sample from its neighbours without itself
        (batch.rs:165-190), or with `folds` = k the k-fold form, each sample
        from the samples outside its fold, holes kept whole (batch.rs:197-243)."""
        if self._values is None:
            raise InvalidInput("call fit first")
        self._ready()
        vg = self._variogram._c
        n = len(self._values)
        if folds is None:
            def one(s, rows):
                if self._block is None:
                    out = self._engine.loo(vg, s._c, rows.astype(np.int64), self._kind, self._param,
                                           self._precision)
                else:
                    out = self._engine.block_loo(vg, s._c, self._block, rows.astype(np.int64), self._kind,
                                                 self._param, self._precision)
                return out["value"], out["variance"], out["status"] == _abi.OK
        else:
            k = int(folds)
            if k < 2:
                raise InvalidInput("k must be ≥ 2")
            fold = self._folds(k)
            S = self._samples
            P = np.stack([S["x"], S["y"], S["z"]], axis=1)
            from .engine import Engine

            eng = Engine(self._device)  # holds one fold's training samples at a time

            def one(s, rows):
                value, variance = np.full(len(rows), np.nan), np.full(len(rows), np.nan)
                ok = np.zeros(len(rows), dtype=bool)
                for f in range(k):
                    test = np.nonzero(fold[rows] == f)[0]
                    if len(test) == 0:
                        continue
                    self._upload(eng, np.nonzero(fold != f)[0])
                    T = P[rows[test]]
                    if self._block is None:
                        out = eng.krige(vg, s._c, T[:, 0].copy(), T[:, 1].copy(), T[:, 2].copy(), self._kind,
                                        self._param, self._precision)
                    else:
                        out = eng.block_krige(vg, s._c, self._block, T[:, 0].copy(), T[:, 1].copy(),
                                              T[:, 2].copy(), self._kind, self._param, self._precision)
                    value[test], variance[test] = out["value"], out["variance"]
                    ok[test] = out["status"] == _abi.OK
                return value, variance, ok

            try:
                value, variance, _ = run_passes(n, self._passes, one)
            finally:
                eng.close()
            return CrossValidation(self._values.copy(), value, variance)
        value, variance, _ = run_passes(n, self._passes, one)
        return CrossValidation(self._values.copy(), value, variance)

    def weights(self, targets):
        """weights_many (crates/estimation/src/batch.rs:76-95) with the first
        search: per target, the indices of the fitted samples it is kriged
        from, in the estimator's order, and their kriging weights; None where
        the target is not estimated."""
        if self._values is None:
            raise InvalidInput("call fit first")
        T = coords_of(targets)
        eng = self._ready()
        out = eng.krige_weights(self._variogram._c, self._passes[0]._c, T[:, 0].copy(), T[:, 1].copy(),
                                T[:, 2].copy(), self._kind, self._param, self._block, self._precision)
        res = []
        for t in range(len(T)):
            c = int(out["count"][t])
            res.append(None if out["status"][t] != _abi.OK else
                       (out["idx"][t, :c].astype(np.int64), out["weights"][t, :c].copy()))
        return res


    @property
    def unit(self):
        """Unit of the fitted values (python/boitata/estimation.py:220-223);
        None, as this engine takes no units."""
        return None

    def with_search(self, search):
        """A copy of the estimator, fitted samples included, with `search` (one
        search or passes) in place of its searches
        (python/boitata/estimation.py:357-375)."""
        passes = _searches(search)
        if self._kind == _abi.SIMPLE and any(p._c.calibration == _abi.CALIBRATE_SLOPE for p in passes):
            raise InvalidInput("simple kriging has a slope of regression of 1; calibrate it with target_efficiency")
        est = type(self).__new__(type(self))
        est.__dict__.update(self.__dict__)
        est._passes = passes
        est._engine = None  # a context of its own; the samples upload on first use
        est._uploaded = False
        return est

    def to_parquet(self, path) -> None:
        """Writes the samples as Parquet columns and the parameters as JSON in
        the file metadata (python/boitata/estimation.py:377-384)."""
        from . import persist

        persist.to_parquet(self, path, type(self).__name__)

    @classmethod
    def from_parquet(cls, path, **engine):
        """Reads `to_parquet` output; predictions match the saved estimator's
        bit for bit. Raises InvalidInput if the file holds another class or a
        newer format, FileError if it cannot be read
        (python/boitata/estimation.py:386-401). `engine` takes this engine's
        own options (precision, device)."""
        from . import persist

        return persist.from_parquet(cls, path, **engine)

    def __reduce__(self):
        # The pickle carries the file's JSON and columns (crates/py/src/persist.rs:90-112)
        # and this engine's own options.
        from . import persist

        meta, cols = persist.state(self, type(self).__name__)
        precision = next(k for k, v in _PRECISIONS.items() if v == self._precision)
        return _unpickle, (type(self), meta, cols, {"precision": precision, "device": self._device})


def _unpickle(cls, meta, columns, engine):
    from . import persist

    return persist.from_state(cls, meta, columns, **engine)


class OrdinaryKriging(_Base):
    """Ordinary kriging: the mean is unknown and taken as constant within each
    neighbourhood, so the kriging weights are constrained to add up to one."""

    def __init__(self, variogram: Variogram, search, **engine):
        super().__init__(variogram, search, **engine)


class SimpleKriging(_Base):
    """Kriging with a known global `mean`; a variogram with a structure without
    a sill (power) raises InvalidInput."""

    _kind = _abi.SIMPLE

    def __init__(self, variogram: Variogram, search, *, mean: float = 0.0, **engine):
        super().__init__(variogram, search, **engine)
        if not variogram.has_sill:
            raise InvalidInput("simple kriging needs a variogram with a sill")
        # py/estimation.rs:820-838
        if any(p._c.calibration == _abi.CALIBRATE_SLOPE for p in self._passes):
            raise InvalidInput("simple kriging has a slope of regression of 1; calibrate it with target_efficiency")
        self._param = float(mean)


class IndicatorKriging(_Base):
    """Ordinary kriging applied to the indicator 1[value <= threshold]: each
    estimate is the probability, clamped to [0, 1], that the value at the
    target does not exceed the threshold."""

    _kind = _abi.INDICATOR

    def __init__(self, variogram: Variogram, search, threshold: float, **engine):
        if isinstance(threshold, str):
            raise NotImplementedError("thresholds with units are not supported by this engine")
        super().__init__(variogram, search, **engine)
        self._param = float(threshold)


class BlockKriging(_Base):
    """Ordinary kriging of block averages; targets are block centres of `size`
    along the world axes, discretised `discretization` per axis; (1, 1, 1) is
    point support, nugget included."""

    def __init__(self, variogram: Variogram, search, size, *, discretization=(4, 4, 1), **engine):
        super().__init__(variogram, search, **engine)
        size = [float(s) for s in size]
        disc = tuple(int(k) for k in discretization)
        if len(size) != 3 or not all(math.isfinite(s) and s > 0 for s in size):
            raise InvalidInput("block size must be three finite values > 0")
        if len(disc) != 3 or min(disc) < 1:
            raise InvalidInput("discretization must be at least 1 per axis")
        self._block = _abi.make_block(size, disc)

    def _one_pass(self, s, T):
        out = self._engine.block_krige(self._variogram._c, s._c, self._block, T[:, 0].copy(), T[:, 1].copy(),
                                       T[:, 2].copy(), self._kind, self._param, self._precision)
        return out["value"], out["variance"], out["status"] == _abi.OK


# --------------------------------------------------------------------------
# Multiple indicator kriging (crates/py/src/indicator.rs:116-420,
# crates/estimation/src/indicator.rs).

_INTERPOLATIONS = {"global": _abi.MIK_GLOBAL, "linear": _abi.MIK_LINEAR}
_TAILS = {"power": _abi.TAIL_POWER, "hyperbolic": _abi.TAIL_HYPERBOLIC}
DIAGNOSTIC_COLUMNS = ("n_samples", "pass", "n_holes", "mean_distance", "max_samples_reached", "correction",
                      "n_order_violations")


@dataclass(frozen=True)
class IndicatorSummary:
    """Conditional distributions per target (indicator.rs:101-116): arrays
    with one row per target; per-threshold, per-cutoff and per-quantile
    arrays have one column each; NaN where unestimated."""

    thresholds: list
    cdf: np.ndarray
    correction: np.ndarray
    mean: np.ndarray
    variance: np.ndarray
    cutoffs: list
    probability_above: np.ndarray
    mean_above: np.ndarray
    quantiles: list
    quantile_values: np.ndarray
    diagnostics: DiagnosticsTable | None = None

    @property
    def std(self) -> np.ndarray:
        return np.sqrt(self.variance)


@dataclass(frozen=True)
class IndicatorCrossValidation(CrossValidation):
    """Cross-validation of multiple indicator kriging (python/boitata/estimation.py:150-185):
    `estimate` is the E-type mean, `variance` the conditional variance, `cdf`
    the (samples, thresholds) corrected probabilities and `pit` F*(actual)."""

    thresholds: list
    cdf: np.ndarray
    pit: np.ndarray

    @property
    def brier(self) -> np.ndarray:
        indicator = self.actual[:, None] <= np.asarray(self.thresholds)
        return np.nanmean((self.cdf - indicator) ** 2, axis=0)

    def accuracy(self, p):
        pit = self.pit[~np.isnan(self.pit)]
        half = np.asarray(p, dtype=float)[..., None] / 2
        return np.mean(np.abs(pit - 0.5) <= half, axis=-1)

    @property
    def goodness(self) -> float:
        p = (np.arange(100) + 0.5) / 100
        excess = self.accuracy(p) - p
        return float(1 - np.mean(np.where(excess >= 0, 1, -2) * excess))


class MultipleIndicatorKriging:
    """Multiple indicator kriging (crates/py/src/indicator.rs:129-192): the
    conditional distribution P(Z <= t | data) kriged at each of `thresholds`
    with its own variogram (or one for all), corrected for order relations,
    completed within classes (`interpolation` "global" or "linear") and above
    the last threshold (`upper_tail` ("power", w) or ("hyperbolic", w)), then
    summarised. The search must not be calibrated."""

    def __init__(self, variogram, search, thresholds, *, simple=False, tails=None, interpolation="global",
                 upper_tail=None, precision: str = "fp64", device: int = 0):
        if isinstance(variogram, Variogram):
            vgs = [variogram]
        elif isinstance(variogram, Sequence) and len(variogram) > 0 and all(isinstance(v, Variogram)
                                                                            for v in variogram):
            vgs = list(variogram)
        else:
            raise InvalidInput("variogram must be a Variogram or a sequence of Variogram")
        if interpolation not in _INTERPOLATIONS:
            raise InvalidInput(f"unknown interpolation {interpolation!r}")
        kind, omega = _abi.TAIL_NONE, 1.0
        if upper_tail is not None:
            name, omega = upper_tail
            if name not in _TAILS:
                raise InvalidInput(f"unknown upper tail {name!r}")
            kind, omega = _TAILS[name], float(omega)
        if precision not in _PRECISIONS:
            raise InvalidInput(f"precision is one of {', '.join(_PRECISIONS)}")
        thr = [_number(t, "threshold") for t in thresholds]
        self._thresholds = thr
        self._variograms = vgs
        self._mik = _abi.make_mik(thr, len(vgs), simple=simple, interpolation=_INTERPOLATIONS[interpolation],
                                  upper_tail=kind, omega=omega, tails=tails)
        self._validate()
        self._passes = _searches(search)
        if any(p._c.calibration for p in self._passes):
            raise InvalidInput("MultipleIndicatorKriging takes a search without target_slope or target_efficiency")
        self._precision = _PRECISIONS[precision]
        self._device = device
        self._engine = None
        self._values = None
        self._weights = None
        self._holes = None
        self._xyz = None
        self._uploaded = False

    def _validate(self):
        # indicator.rs:291-317
        t, p = self._thresholds, self._mik
        if not t or not all(math.isfinite(v) for v in t):
            raise InvalidInput("need at least one finite threshold")
        if any(a >= b for a, b in zip(t, t[1:])):
            raise InvalidInput("thresholds must be strictly increasing")
        if len(self._variograms) not in (1, len(t)):
            raise InvalidInput("need one variogram, or one per threshold")
        if p.has_tails and not (math.isfinite(p.tails[0]) and math.isfinite(p.tails[1]) and p.tails[0] <= p.tails[1]):
            raise InvalidInput("tails must be finite with lower <= upper")
        if p.upper_tail != _abi.TAIL_NONE and not (math.isfinite(p.omega) and p.omega > 0.0):
            raise InvalidInput("the upper tail exponent must be > 0")
        if p.upper_tail == _abi.TAIL_HYPERBOLIC and t[-1] <= 0.0:
            raise InvalidInput("a hyperbolic upper tail needs a last threshold > 0")

    @property
    def thresholds(self):
        return list(self._thresholds)

    def _vgs(self):
        return [v._c for v in self._variograms]

    def fit(self, coords, values, *, weights=None, holes=None):
        """Stores the samples; samples sharing a location keep the first one.
        `weights` decluster the global distribution (the simple kriging means,
        the class contents and the tails)."""
        if isinstance(values, str) or isinstance(weights, str):
            raise NotImplementedError("columns by name need a container; pass arrays")
        xyz = coords_of(coords)
        n = len(xyz)
        v = np.asarray(values, dtype=np.float64).reshape(-1)
        if len(v) != n:
            raise InvalidInput(f"values: expected {n} values, got {len(v)}")
        if not np.isfinite(v).all():
            raise InvalidInput("values must be finite")
        w = None
        if weights is not None:
            w = np.asarray(weights, dtype=np.float64).reshape(-1)
            if len(w) != n:
                raise InvalidInput(f"weights: expected {n} values, got {len(w)}")
            if not np.isfinite(w).all() or (w < 0).any() or w.sum() <= 0:
                raise InvalidInput("weights must be >= 0 and not all 0")
        text, ids = hole_codes(holes, n)
        keep, note = distinct_rows(xyz, text)
        if note:
            warnings.warn(note, UserWarning, stacklevel=2)
        xyz, v = xyz[keep], v[keep]
        w = None if w is None else w[keep]
        ids = None if ids is None else ids[keep]
        self._values, self._weights, self._holes, self._xyz = v, w, ids, np.ascontiguousarray(xyz)
        self._uploaded = False
        self._fitted()
        return self

    def _fitted(self):
        """Raises before fit; uploads the fitted samples on first use (after a
        restore from a file or a pickle)."""
        if self._values is None:
            raise InvalidInput("call fit first")
        if not self._uploaded:
            if self._engine is None:
                from .engine import Engine

                self._engine = Engine(self._device)
            x = self._xyz
            self._engine.set_samples(x[:, 0].copy(), x[:, 1].copy(), x[:, 2].copy(), self._values, None, self._holes)
            self._uploaded = True

    def to_parquet(self, path) -> None:
        """Writes the samples (and declustering weights) as Parquet columns and
        the parameters as JSON in the file metadata (crates/py/src/indicator.rs:505-525)."""
        from . import persist

        persist.to_parquet(self, path, type(self).__name__)

    @classmethod
    def from_parquet(cls, path, **engine):
        """Reads `to_parquet` output; predictions match the saved estimator's
        bit for bit. `engine` takes this engine's own options (precision,
        device)."""
        from . import persist

        return persist.from_parquet(cls, path, **engine)

    def __reduce__(self):
        from . import persist

        meta, cols = persist.state(self, type(self).__name__)
        precision = next(k for k, v in _PRECISIONS.items() if v == self._precision)
        return _unpickle, (type(self), meta, cols, {"precision": precision, "device": self._device})

    def predict(self, targets, *, cutoffs=None, quantiles=(), anisotropy=None, diagnostics=False,
                discretization=None, block_size=None) -> IndicatorSummary:
        """Conditional distribution at each target, through the search passes.
        With `discretization` the indicators are kriged over blocks of
        `block_size` centred on the targets (the reference takes both from a
        BlockModel)."""
        if anisotropy is not None:
            raise NotImplementedError("local anisotropy is not supported by this engine")
        self._fitted()
        cut = [] if cutoffs is None else [_number(c, "cutoff") for c in np.atleast_1d(cutoffs)]
        qs = [float(q) for q in quantiles]
        block = None
        if discretization is not None:
            if block_size is None:
                raise InvalidInput("discretization needs BlockModel targets: give block_size")
            block = _abi.make_block([float(s) for s in block_size], tuple(int(k) for k in discretization))
        T = coords_of(targets)
        m, K, C, Q = len(T), len(self._thresholds), len(cut), len(qs)
        res = dict(cdf=np.full((K, m), np.nan), correction=np.full(m, np.nan), mean=np.full(m, np.nan),
                   variance=np.full(m, np.nan), probability_above=np.full((C, m), np.nan),
                   mean_above=np.full((C, m), np.nan), quantile_values=np.full((Q, m), np.nan),
                   violations=np.zeros(m), n_used=np.zeros(m))
        which = np.full(m, -1, dtype=np.int64)
        remaining = np.arange(m)
        for p, s in enumerate(self._passes):
            if len(remaining) == 0:
                break
            R = T[remaining]
            out = self._engine.mik_predict(self._mik, self._thresholds, self._vgs(), s._c, R[:, 0].copy(),
                                           R[:, 1].copy(), R[:, 2].copy(), weights=self._weights, block=block,
                                           cutoffs=cut, quantiles=qs, precision=self._precision)
            ok = out["status"] == _abi.OK
            rows = remaining[ok]
            for k in res:
                src = out[k]
                res[k][..., rows] = src[..., ok]
            which[rows] = p
            remaining = remaining[~ok]
        diag = self._diagnostics(T, which, res) if diagnostics else None
        return IndicatorSummary(list(self._thresholds), res["cdf"].T.copy(), res["correction"], res["mean"],
                                res["variance"], cut, res["probability_above"].T.copy(), res["mean_above"].T.copy(),
                                qs, res["quantile_values"].T.copy(), diag)

    def _diagnostics(self, T, which, res):
        """indicator.rs:793-816 with neighborhood_stats (neighborhood.rs:66-136)
        over each target's samples: their count, distinct holes (untagged
        samples count one each), mean Euclidean distance, in ascending order."""
        m = len(T)
        n_holes = np.full(m, np.nan)
        mean_d = np.full(m, np.nan)
        reached = np.full(m, np.nan)
        svg = self._variograms[len(self._variograms) // 2]._c
        for p, s in enumerate(self._passes):
            rows = np.nonzero(which == p)[0]
            if len(rows) == 0:
                continue
            R = T[rows]
            idx, cnt, _ = self._engine.neighbors(s._c, svg, R[:, 0].copy(), R[:, 1].copy(), R[:, 2].copy())
            for r, t in enumerate(rows):
                sel = idx[r, : cnt[r]]
                d = self._xyz[sel] - T[t]
                dist = np.sort(np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]))
                mean_d[t] = np.cumsum(dist)[-1] / len(dist)
                n_holes[t] = len(sel) if self._holes is None else len(np.unique(self._holes[sel]))
                reached[t] = 1.0 if cnt[r] >= s.max_samples else 0.0
        ok = which >= 0
        n_samples = np.where(ok, res["n_used"], np.nan)
        viol = np.where(ok, res["violations"], np.nan)
        passes = np.where(ok, which + 1.0, np.nan)
        return DiagnosticsTable(zip(DIAGNOSTIC_COLUMNS, (n_samples, passes, n_holes, mean_d, reached,
                                                         res["correction"], viol)))

    def cross_validate(self, *, folds=None) -> IndicatorCrossValidation:
        """Re-estimates each sample's distribution from the others, through the
        same search passes, against the global distribution of all samples."""
        if folds is not None:
            raise NotImplementedError("k-fold cross-validation is not supported by this engine")
        self._fitted()
        n, K = len(self._values), len(self._thresholds)
        cdf = np.full((K, n), np.nan)
        mean, var, pit = np.full(n, np.nan), np.full(n, np.nan), np.full(n, np.nan)
        remaining = np.arange(n)
        for s in self._passes:
            if len(remaining) == 0:
                break
            out = self._engine.mik_cross_validate(self._mik, self._thresholds, self._vgs(), s._c,
                                                  weights=self._weights, which=remaining.astype(np.int64),
                                                  precision=self._precision)
            ok = out["status"] == _abi.OK
            rows = remaining[ok]
            cdf[:, rows] = out["cdf"][:, ok]
            mean[rows], var[rows], pit[rows] = out["mean"][ok], out["variance"][ok], out["pit"][ok]
            remaining = remaining[~ok]
        return IndicatorCrossValidation(self._values.copy(), mean, var, list(self._thresholds), cdf.T.copy(), pit)
