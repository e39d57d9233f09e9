"""Persistence of the kriging estimators, searches and variograms, in the
reference's formats (crates/py/src/persist.rs):

* JSON envelopes ``{"type": <class>, "format": 1, <extra>, <fields>...}``
  (persist.rs:19-56): the fields are the serde layout of the reference's
  types, restated here field for field so that files move both ways —
  ``Search`` (crates/estimation/src/search.rs:30-81 flattened into
  crates/py/src/estimation.rs:210-221), ``Variogram``
  (crates/variogram/src/composite.rs:12-18, model.rs:10-38, aniso.rs:17-27),
  and the estimator (crates/py/src/estimation.rs:696-725,1213-1231).
* Floats that may be NaN or infinite (the search radius, the calibration
  targets) are written as the strings "NaN", "inf" and "-inf"
  (crates/core/src/nonfinite.rs).
* A fitted estimator is a Parquet file: the samples as float64 columns x, y,
  z, value, hole, error_variance and domain (null when not given), the
  parameters as JSON under the schema metadata key ``boitata``, with
  ``"fitted": true`` (persist.rs:90-133, crates/py/src/estimation.rs:1110-1121,
  crates/io/src/parquet.rs:487-497).
* A pickle carries the same JSON and columns (persist.rs:90-112).

Multiple indicator kriging is written as crates/py/src/indicator.rs:115-125
lays it out (the MultipleIndicator model, crates/estimation/src/indicator.rs:
34-65, and the passes), its declustering weights as a ``weight`` column.

Extensions of this engine with no reference field are written only when
used: a structure's own ``ratios`` (the per-structure anisotropy ratios).
"""
from __future__ import annotations

import json
import math
import os

import numpy as np

from .errors import FileError, InvalidInput

FORMAT = 1
KEY = "boitata"
SAMPLE_COLUMNS = ("x", "y", "z", "value", "hole", "error_variance", "domain")

# crates/variogram/src/model.rs:10-24: unit variants are strings, the two with
# a parameter are one-key objects.
_MODEL_NAMES = {"spherical": "Spherical", "exponential": "Exponential", "gaussian": "Gaussian", "cubic": "Cubic",
                "pentaspherical": "PentaSpherical", "circular": "Circular", "sinehole": "SineHole"}
_MODEL_KEYS = {v: k for k, v in _MODEL_NAMES.items()}


def _float(v: float):
    """crates/core/src/nonfinite.rs:21-31."""
    v = float(v)
    if math.isfinite(v):
        return v
    if math.isnan(v):
        return "NaN"
    return "inf" if v > 0 else "-inf"


def _parse_float(v, what):
    """crates/core/src/nonfinite.rs:33-43."""
    if isinstance(v, bool) or v is None:
        raise InvalidInput(f"{what}: expected a number")
    if isinstance(v, (int, float)):
        return float(v)
    if isinstance(v, str):
        table = {"NaN": math.nan, "inf": math.inf, "-inf": -math.inf}
        if v in table:
            return table[v]
        raise InvalidInput(f"{what}: {v!r} is not a number")
    raise InvalidInput(f"{what}: expected a number")


def _number(fields, key, what=None):
    v = fields.get(key)
    if isinstance(v, bool) or not isinstance(v, (int, float)):
        raise InvalidInput(f"{what or key}: expected a number")
    return float(v)


def _count(fields, key, optional=False):
    v = fields.get(key)
    if v is None and optional:
        return None
    if isinstance(v, bool) or not isinstance(v, int) or v < 0:
        raise InvalidInput(f"{key}: expected a non-negative integer")
    return v


def envelope(name: str, fields: dict, extra: dict | None = None) -> str:
    """persist.rs:21-32: type and format first, then `extra`, then the fields."""
    obj = {"type": name, "format": FORMAT}
    obj.update(extra or {})
    obj.update(fields)
    return json.dumps(obj, separators=(",", ":"), allow_nan=False)


def open_envelope(name: str, text: str) -> dict:
    """persist.rs:35-51: the fields once the type and format are checked."""
    try:
        obj = json.loads(text)
    except (TypeError, ValueError) as e:
        raise InvalidInput(str(e)) from None
    if not isinstance(obj, dict):
        raise InvalidInput("expected a JSON object")
    found = obj.pop("type", None)
    if found != name:
        raise InvalidInput(f"expected a {name}, found {json.dumps(found)}")
    f = obj.pop("format", None)
    if isinstance(f, bool) or not isinstance(f, int) or f < 0:
        raise InvalidInput("missing or invalid format number")
    if f > FORMAT:
        raise InvalidInput(f"format {f} is newer than this boitata reads (up to {FORMAT}); upgrade boitata")
    return obj


# --- anisotropy (crates/variogram/src/aniso.rs:17-27, crates/py/src/variogram.rs:164-177)

def _angles(rotation, ratios):
    """None for the identity (rotation 0, ratios 1), else the Angles object
    with the major range 1."""
    rot = tuple(float(a) for a in rotation)
    rat = tuple(float(r) for r in ratios)
    if rot == (0.0, 0.0, 0.0) and rat == (1.0, 1.0):
        return None
    return {"azimuth": rot[0], "dip": rot[1], "rake": rot[2], "major": 1.0, "semi": rat[0], "minor": rat[1]}


def _from_angles(a):
    """(rotation, ratios) of an Angles object (`pitch` is an alias of `rake`)."""
    if a is None:
        return None, None
    if not isinstance(a, dict):
        raise InvalidInput("anisotropy: expected an object")
    rake = a.get("rake", a.get("pitch"))
    rot = (_number(a, "azimuth"), _number(a, "dip"), _parse_float(rake, "rake") if rake is not None else None)
    if rot[2] is None:
        raise InvalidInput("anisotropy: missing rake")
    major, semi, minor = (_number(a, k) for k in ("major", "semi", "minor"))
    if not all(r > 0.0 and math.isfinite(r) for r in (major, semi, minor)):
        raise InvalidInput("all ranges must be finite and > 0")
    if major != 1.0:
        raise InvalidInput("an anisotropy whose major range is not 1 is not supported by this engine")
    return rot, (semi, minor)


# --- variogram ----------------------------------------------------------------

def variogram_fields(vg) -> dict:
    """composite.rs:12-18: nugget, structures, anisotropy."""
    structures = []
    for s in vg.structures:
        if s.model == "matern":
            model = {"Matern": {"order": float(s.param)}}
        elif s.model == "power":
            model = {"Power": {"exponent": float(s.param)}}
        else:
            model = _MODEL_NAMES[s.model]
        entry = {"model": model, "sill": float(s.sill), "range": float(s.range)}
        if s.ratios is not None:
            entry["ratios"] = [float(s.ratios[0]), float(s.ratios[1])]
        structures.append(entry)
    return {"nugget": float(vg.nugget), "structures": structures, "anisotropy": _angles(vg.rotation, vg.ratios)}


def variogram_from(fields: dict):
    from .estimation import Structure, Variogram

    if not isinstance(fields, dict):
        raise InvalidInput("variogram: expected an object")
    if fields.get("length_unit") is not None or fields.get("unit") is not None:
        raise NotImplementedError("length units are not supported by this engine")
    structures = []
    for e in fields.get("structures") or []:
        if not isinstance(e, dict):
            raise InvalidInput("structures: expected objects")
        m = e.get("model")
        kw = {}
        if isinstance(m, str) and m in _MODEL_KEYS:
            name = _MODEL_KEYS[m]
        elif isinstance(m, dict) and list(m) == ["Matern"] and isinstance(m["Matern"], dict):
            name, kw = "matern", {"order": _number(m["Matern"], "order")}
        elif isinstance(m, dict) and list(m) == ["Power"] and isinstance(m["Power"], dict):
            name, kw = "power", {"exponent": _number(m["Power"], "exponent")}
        else:
            raise InvalidInput(f"unknown variogram model {json.dumps(m)}")
        if e.get("ratios") is not None:
            kw["ratios"] = tuple(float(r) for r in e["ratios"])
        structures.append(Structure(name, _number(e, "sill"), _number(e, "range"), **kw))
    rot, rat = _from_angles(fields.get("anisotropy"))
    return Variogram(structures, nugget=_number(fields, "nugget"), rotation=rot or (0.0, 0.0, 0.0),
                     ratios=rat or (1.0, 1.0))


# --- search -----------------------------------------------------------------

def search_fields(s) -> dict:
    """search.rs:30-81 in declaration order; optional fields marked
    skip_serializing_if are left out when unset, the others written null."""
    from . import _abi

    c = s._c
    out = {"min_samples": int(c.min_samples), "max_samples": int(c.max_samples), "radius": _float(c.radius),
           "max_per_hole": None if c.max_per_hole < 0 else int(c.max_per_hole), "octant": bool(c.octant)}
    if c.sectors > 0:
        out["sectors"] = {"count": int(c.sectors), "max_per_sector": int(c.max_per_sector)}
    out["anisotropy"] = (_angles((c.azimuth, c.dip, c.rake), (c.ratio_semi, c.ratio_minor))
                         if c.has_anisotropy else None)
    out["high_grade"] = None
    if c.calibration == _abi.CALIBRATE_SLOPE:
        out["calibration"] = {"slope": _float(c.calibration_target)}
    elif c.calibration == _abi.CALIBRATE_EFFICIENCY:
        out["calibration"] = {"efficiency": _float(c.calibration_target)}
    for k in ("max_per_octant", "min_sectors", "max_empty_sectors", "min_holes"):
        v = getattr(c, k)
        if v >= 0:
            out[k] = int(v)
    return out


def search_from(fields: dict):
    from .estimation import Search

    if not isinstance(fields, dict):
        raise InvalidInput("search: expected an object")
    for k in ("high_grade", "soft", "length_unit"):
        if fields.get(k) is not None:
            raise NotImplementedError(f"{k} is not supported by this engine")
    kw = {}
    sectors = fields.get("sectors")
    if sectors is not None:
        if not isinstance(sectors, dict):
            raise InvalidInput("sectors: expected an object")
        kw["sectors"], kw["max_per_sector"] = _count(sectors, "count"), _count(sectors, "max_per_sector")
    per_octant = _count(fields, "max_per_octant", optional=True)
    octant = fields.get("octant", False)
    if not isinstance(octant, bool):
        raise InvalidInput("octant: expected true or false")
    if per_octant is not None:
        if not octant or sectors is not None:
            raise InvalidInput("max_per_octant needs octant without sectors")
        kw["max_per_sector"] = per_octant
    cal = fields.get("calibration")
    if cal is not None:
        if not isinstance(cal, dict) or len(cal) != 1 or next(iter(cal)) not in ("slope", "efficiency"):
            raise InvalidInput("calibration: expected {\"slope\": x} or {\"efficiency\": x}")
        k, v = next(iter(cal.items()))
        kw["target_" + k] = _parse_float(v, "calibration")
    rot, rat = _from_angles(fields.get("anisotropy"))
    if rot is not None:
        kw["rotation"], kw["ratios"] = rot, rat
    return Search(_parse_float(fields.get("radius"), "radius"), max_samples=_count(fields, "max_samples"),
                  min_samples=_count(fields, "min_samples"), octant=octant,
                  max_per_hole=_count(fields, "max_per_hole", optional=True),
                  min_sectors=_count(fields, "min_sectors", optional=True),
                  max_empty_sectors=_count(fields, "max_empty_sectors", optional=True),
                  min_holes=_count(fields, "min_holes", optional=True), **kw)


# --- estimators ---------------------------------------------------------------

def method_fields(est) -> dict:
    """crates/py/src/estimation.rs:696-725 (externally tagged enums): the
    kriging kind, or block support with its size and discretisation."""
    from .estimation import BlockKriging, IndicatorKriging, SimpleKriging

    if isinstance(est, BlockKriging):
        b = est._block
        return {"Block": {"size": [float(b.size[d]) for d in range(3)],
                          "disc": {"nx": int(b.nx), "ny": int(b.ny), "nz": int(b.nz)},
                          "rotation": [float(b.rotation[d]) for d in range(3)]}}
    if isinstance(est, SimpleKriging):
        return {"Kriging": {"Simple": {"mean": float(est._param)}}}
    if isinstance(est, IndicatorKriging):
        return {"Kriging": {"Indicator": {"threshold": float(est._param)}}}
    return {"Kriging": "Ordinary"}


def estimator_fields(est) -> dict:
    """crates/py/src/estimation.rs:1213-1231: method, variogram, the search
    passes, drift data (none) and the (value, coordinate, length) units."""
    return {"method": method_fields(est), "variogram": variogram_fields(est._variogram),
            "search": [search_fields(s) for s in est._passes], "drift_data": [], "units": [None, None, None]}


def estimator_from(cls, fields: dict, **engine):
    """An unfitted estimator of class `cls` from its fields; the method must be
    the class's own."""
    from .estimation import BlockKriging, IndicatorKriging, OrdinaryKriging, SimpleKriging

    if fields.get("domains") is not None:
        raise NotImplementedError("domains are not supported by this engine")
    if fields.get("drift_data"):
        raise NotImplementedError("external drift is not supported by this engine")
    units = fields.get("units") or [None, None, None]
    if any(u is not None for u in units):
        raise NotImplementedError("units are not supported by this engine")
    if fields.get("variogram") is None:
        raise InvalidInput("kriging needs a variogram")
    vg = variogram_from(fields["variogram"])
    passes = fields.get("search")
    if not isinstance(passes, list) or not passes:
        raise InvalidInput("search: expected a list of searches")
    search = [search_from(s) for s in passes]
    m = fields.get("method")
    if cls is BlockKriging and isinstance(m, dict) and list(m) == ["Block"] and isinstance(m["Block"], dict):
        b = m["Block"]
        d = b.get("disc")
        if not isinstance(d, dict):
            raise InvalidInput("block: missing discretization")
        if any(float(r) != 0.0 for r in b.get("rotation", [0.0, 0.0, 0.0])):
            raise NotImplementedError("rotated blocks are not supported by this engine")
        return BlockKriging(vg, search, [float(v) for v in b.get("size", ())],
                            discretization=(_count(d, "nx"), _count(d, "ny"), _count(d, "nz")), **engine)
    k = m.get("Kriging") if isinstance(m, dict) and list(m) == ["Kriging"] else None
    if cls is OrdinaryKriging and k == "Ordinary":
        return OrdinaryKriging(vg, search, **engine)
    if cls is SimpleKriging and isinstance(k, dict) and list(k) == ["Simple"]:
        return SimpleKriging(vg, search, mean=_number(k["Simple"], "mean"), **engine)
    if cls is IndicatorKriging and isinstance(k, dict) and list(k) == ["Indicator"]:
        return IndicatorKriging(vg, search, _number(k["Indicator"], "threshold"), **engine)
    raise InvalidInput(f"method {json.dumps(m)} is not that of a {cls.__name__}")


# --- multiple indicator kriging (crates/py/src/indicator.rs:115-125,505-525;
# crates/estimation/src/indicator.rs:34-65)

_INTERP_NAMES = {"global": "Global", "linear": "Linear"}
_TAIL_NAMES = {"power": "Power", "hyperbolic": "Hyperbolic"}


def mik_fields(est) -> dict:
    """The MultipleIndicator model and the search passes."""
    from . import _abi
    from .estimation import _INTERPOLATIONS, _TAILS

    p = est._mik
    interp = next(k for k, v in _INTERPOLATIONS.items() if v == p.interpolation)
    tail = None
    if p.upper_tail != _abi.TAIL_NONE:
        name = next(k for k, v in _TAILS.items() if v == p.upper_tail)
        tail = {_TAIL_NAMES[name]: float(p.omega)}
    model = {"thresholds": [float(t) for t in est._thresholds],
             "variograms": [variogram_fields(v) for v in est._variograms], "simple": bool(p.simple),
             "tails": [float(p.tails[0]), float(p.tails[1])] if p.has_tails else None,
             "interpolation": _INTERP_NAMES[interp], "upper_tail": tail}
    return {"model": model, "search": [search_fields(s) for s in est._passes]}


def mik_from(cls, fields: dict, **engine):
    if fields.get("unit") is not None:
        raise NotImplementedError("units are not supported by this engine")
    m = fields.get("model")
    if not isinstance(m, dict):
        raise InvalidInput("model: expected an object")
    thr = m.get("thresholds")
    if not isinstance(thr, list):
        raise InvalidInput("thresholds: expected a list")
    vgs = m.get("variograms")
    if not isinstance(vgs, list) or not vgs:
        raise InvalidInput("variograms: expected a list")
    interp = {v: k for k, v in _INTERP_NAMES.items()}.get(m.get("interpolation"))
    if interp is None:
        raise InvalidInput(f"unknown interpolation {json.dumps(m.get('interpolation'))}")
    tail = m.get("upper_tail")
    upper = None
    if tail is not None:
        names = {v: k for k, v in _TAIL_NAMES.items()}
        if not isinstance(tail, dict) or len(tail) != 1 or next(iter(tail)) not in names:
            raise InvalidInput(f"unknown upper tail {json.dumps(tail)}")
        k, w = next(iter(tail.items()))
        upper = (names[k], _parse_float(w, "upper_tail"))
    tails = m.get("tails")
    if tails is not None:
        if not isinstance(tails, list) or len(tails) != 2:
            raise InvalidInput("tails: expected a pair")
        tails = (_parse_float(tails[0], "tails"), _parse_float(tails[1], "tails"))
    simple = m.get("simple", False)
    if not isinstance(simple, bool):
        raise InvalidInput("simple: expected true or false")
    passes = fields.get("search")
    if not isinstance(passes, list) or not passes:
        raise InvalidInput("search: expected a list of searches")
    return cls([variogram_from(v) for v in vgs], [search_from(s) for s in passes],
               [_parse_float(t, "thresholds") for t in thr], simple=simple, tails=tails, interpolation=interp,
               upper_tail=upper, **engine)


def _is_mik(est) -> bool:
    return hasattr(est, "_mik")


def sample_columns(est) -> dict | None:
    """crates/py/src/estimation.rs:1110-1121 (and the declustering weights of
    multiple indicator kriging, crates/py/src/indicator.rs:510-518); None
    before fit."""
    if _is_mik(est):
        if est._values is None:
            return None
        n = len(est._values)
        x = est._xyz
        cols = {"x": x[:, 0].copy(), "y": x[:, 1].copy(), "z": x[:, 2].copy(), "value": est._values,
                "hole": np.full(n, np.nan) if est._holes is None else est._holes.astype(np.float64),
                "error_variance": np.zeros(n), "domain": np.full(n, np.nan)}
        if est._weights is not None:
            cols["weight"] = est._weights
        return cols
    s = est._samples
    if s is None:
        return None
    n = len(s["value"])
    return {"x": s["x"], "y": s["y"], "z": s["z"], "value": s["value"], "hole": s["hole"],
            "error_variance": s["error_variance"], "domain": np.full(n, np.nan)}


def _values(columns, name, nullable=False):
    if name not in columns:
        raise InvalidInput(f"missing column {name}")
    v = np.ascontiguousarray(columns[name], dtype=np.float64)
    if not nullable and np.isnan(v).any():
        raise InvalidInput(f"null in column {name}")
    return v


def restore(est, columns: dict) -> None:
    """crates/py/src/estimation.rs:1123-1160: the fitted samples from their
    columns (hole codes are indices, null without holes)."""
    x, y, z, v = (_values(columns, k) for k in ("x", "y", "z", "value"))
    hole, err = _values(columns, "hole", nullable=True), _values(columns, "error_variance")
    dom = _values(columns, "domain", nullable=True)
    n = len(x)
    for name, a in (("y", y), ("z", z), ("value", v), ("hole", hole), ("error_variance", err), ("domain", dom)):
        if len(a) != n:
            raise InvalidInput(f"{name}: expected {n} values, got {len(a)}")
    tagged = hole[~np.isnan(hole)]
    if ((tagged < 0) | (tagged != np.floor(tagged)) | (tagged >= 2**31 - 1)).any():
        raise InvalidInput("hole codes must be indices")
    if not np.isnan(dom).all():
        raise NotImplementedError("domains are not supported by this engine")
    if _is_mik(est):
        if (err != 0.0).any():
            raise InvalidInput("multiple indicator kriging takes no error_variance")
        ids = None
        if not np.isnan(hole).all():
            # A sample without a code counts as a hole of its own.
            ids = np.empty(n, dtype=np.int32)
            tagged = ~np.isnan(hole)
            ids[tagged] = hole[tagged].astype(np.int32)
            top = int(hole[tagged].max()) + 1 if tagged.any() else 0
            ids[~tagged] = top + np.arange(int((~tagged).sum()), dtype=np.int32)
        w = _values(columns, "weight") if "weight" in columns else None
        if w is not None and (len(w) != n or (w < 0).any() or not w.sum() > 0):
            raise InvalidInput("weights must be >= 0 and not all 0")
        est._values, est._weights, est._holes = v, w, ids
        est._xyz = np.ascontiguousarray(np.stack([x, y, z], 1))
    else:
        est._samples = dict(x=x, y=y, z=z, value=v, hole=hole, error_variance=err)
        est._values = v
    est._uploaded = False


def state(est, name: str):
    """persist.rs:90-95: the JSON with `fitted`, and the columns (None unfitted)."""
    cols = sample_columns(est)
    fields = mik_fields(est) if _is_mik(est) else estimator_fields(est)
    return envelope(name, fields, {"fitted": cols is not None}), cols


def from_state(cls, meta: str, columns, **engine):
    """persist.rs:97-112."""
    if columns is not None:
        lengths = {len(np.asarray(c)) for c in columns.values()}
        if len(lengths) > 1:
            raise InvalidInput("columns differ in length")
    fields = open_envelope(cls.__name__, meta)
    fitted = fields.pop("fitted", None) is True
    est = mik_from(cls, fields, **engine) if hasattr(cls, "_validate") else estimator_from(cls, fields, **engine)
    if fitted and columns is not None:
        restore(est, columns)
    return est


def to_parquet(est, path, name: str) -> None:
    """persist.rs:114-133: the columns as nullable float64, the JSON under the
    `boitata` schema metadata key."""
    import pyarrow as pa
    import pyarrow.parquet as pq

    meta, cols = state(est, name)
    cols = cols or {k: np.zeros(0) for k in SAMPLE_COLUMNS}
    names = list(cols)
    arrays = [pa.array(np.asarray(cols[k], dtype=np.float64), type=pa.float64(),
                       mask=np.isnan(np.asarray(cols[k], dtype=np.float64)) if k in ("hole", "domain") else None)
              for k in names]
    schema = pa.schema([pa.field(k, pa.float64(), nullable=True) for k in names], metadata={KEY: meta})
    try:
        pq.write_table(pa.Table.from_arrays(arrays, schema=schema), os.fspath(path))
    except OSError as e:
        raise FileError(str(e)) from None


def from_parquet(cls, path, **engine):
    """persist.rs:135-151."""
    import pyarrow as pa
    import pyarrow.parquet as pq

    try:
        table = pq.read_table(os.fspath(path))
    except FileNotFoundError as e:
        raise FileError(str(e)) from None
    except (OSError, pa.ArrowInvalid) as e:
        if not os.path.exists(os.fspath(path)):
            raise FileError(str(e)) from None
        raise InvalidInput(str(e)) from None
    meta = (table.schema.metadata or {}).get(KEY.encode())
    if meta is None:
        raise InvalidInput("no boitata metadata")
    columns = {}
    for f, col in zip(table.schema, table.columns):
        if f.type != pa.float64():
            raise InvalidInput(f"column {f.name} is not float64")
        columns[f.name] = col.to_numpy(zero_copy_only=False).astype(np.float64)
    return from_state(cls, meta.decode(), columns, **engine)
