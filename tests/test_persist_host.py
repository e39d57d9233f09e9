"""Persistence on the host (crates/py/src/persist.rs): the JSON envelopes of
Search and Variogram in the reference's serde layouts, written out by hand
here from the type declarations (crates/estimation/src/search.rs:30-81,
crates/py/src/estimation.rs:210-221, crates/variogram/src/composite.rs:12-18,
model.rs:10-38, aniso.rs:17-27, crates/core/src/nonfinite.rs), the envelope's
type and format checks (persist.rs:35-51), and estimator files laid out as
the reference writes them (a Parquet table of samples with the JSON under
the `boitata` metadata key, crates/py/src/estimation.rs:1110-1121) read back
and written again unchanged. Predictions of reloaded estimators are compared
on the device (test_gpu_persist.py)."""
import json
import pickle

import numpy as np
import pyarrow as pa
import pyarrow.parquet as pq
import pytest

from boitata_amd.errors import FileError, InvalidInput
from boitata_amd.estimation import (BlockKriging, IndicatorKriging, OrdinaryKriging, Search, SimpleKriging,
                                    Structure, Variogram)

ANGLES = '{"azimuth":30.0,"dip":20.0,"rake":0.0,"major":1.0,"semi":0.625,"minor":0.2}'


@pytest.mark.parametrize("search,text", [
    # Options left unset: max_per_hole, anisotropy and high_grade are written
    # null, every field marked skip_serializing_if is left out.
    (Search(50.0),
     '{"type":"Search","format":1,"min_samples":1,"max_samples":16,"radius":50.0,"max_per_hole":null,'
     '"octant":false,"anisotropy":null,"high_grade":null}'),
    # ranges -> radius = major, ratios = semi/major, minor/major; octant caps
    # under max_per_octant; calibration as a one-key object.
    (Search(ranges=(400, 250, 80), rotation=(30, 20, 0), max_samples=32, min_samples=4, octant=True,
            max_per_sector=4, max_per_hole=3, target_slope=0.9),
     '{"type":"Search","format":1,"min_samples":4,"max_samples":32,"radius":400.0,"max_per_hole":3,'
     '"octant":true,"anisotropy":' + ANGLES + ',"high_grade":null,"calibration":{"slope":0.9},'
     '"max_per_octant":4}'),
    # Plane sectors with the default cap ceil(max_samples / count); the sector
    # rules; an infinite radius as a string.
    (Search(float("inf"), max_samples=10, sectors=4, min_sectors=2, max_empty_sectors=1, min_holes=3,
            target_efficiency=0.25),
     '{"type":"Search","format":1,"min_samples":1,"max_samples":10,"radius":"inf","max_per_hole":null,'
     '"octant":false,"sectors":{"count":4,"max_per_sector":3},"anisotropy":null,"high_grade":null,'
     '"calibration":{"efficiency":0.25},"min_sectors":2,"max_empty_sectors":1,"min_holes":3}'),
])
def test_search_json_layout(search, text):
    assert search.to_json() == text
    back = Search.from_json(text)
    assert back.to_json() == text
    assert bytes(back._c) == bytes(search._c)
    assert pickle.loads(pickle.dumps(search)).to_json() == text


@pytest.mark.parametrize("vg,text", [
    (Variogram([("spherical", 1.0, 30.0)]),
     '{"type":"Variogram","format":1,"nugget":0.0,"structures":[{"model":"Spherical","sill":1.0,"range":30.0}],'
     '"anisotropy":null}'),
    (Variogram([("exponential", 0.5, 150.0), ("pentaspherical", 0.2, 60.0), Structure("matern", 0.1, 40.0, order=1.5),
                ("sinehole", 0.1, 20.0)], nugget=0.1, rotation=(30, 20, 0), ratios=(0.625, 0.2)),
     '{"type":"Variogram","format":1,"nugget":0.1,"structures":[{"model":"Exponential","sill":0.5,"range":150.0},'
     '{"model":"PentaSpherical","sill":0.2,"range":60.0},{"model":{"Matern":{"order":1.5}},"sill":0.1,'
     '"range":40.0},{"model":"SineHole","sill":0.1,"range":20.0}],"anisotropy":' + ANGLES + '}'),
    (Variogram([Structure("power", 0.5, 1.0, exponent=1.5)], nugget=0.2),
     '{"type":"Variogram","format":1,"nugget":0.2,"structures":[{"model":{"Power":{"exponent":1.5}},"sill":0.5,'
     '"range":1.0}],"anisotropy":null}'),
])
def test_variogram_json_layout(vg, text):
    assert vg.to_json() == text
    back = Variogram.from_json(text)
    assert back.to_json() == text
    assert bytes(back._c) == bytes(vg._c)
    assert pickle.loads(pickle.dumps(vg)).to_json() == text


def test_reads_what_serde_also_accepts():
    # `pitch` is an alias of `rake`; integral numbers; fields in another order.
    text = ('{"anisotropy":{"azimuth":30,"dip":20,"pitch":0,"major":1,"semi":0.625,"minor":0.2},"format":1,'
            '"nugget":0,"structures":[{"range":30,"sill":1,"model":"Gaussian"}],"type":"Variogram"}')
    vg = Variogram.from_json(text)
    assert vg.to_json() == Variogram([("gaussian", 1.0, 30.0)], rotation=(30, 20, 0), ratios=(0.625, 0.2)).to_json()


@pytest.mark.parametrize("text,msg", [
    ('{"type":"Search","format":1}', "expected a Variogram, found \"Search\""),
    ('{"format":1}', "expected a Variogram, found null"),
    ('{"type":"Variogram","format":2,"nugget":0.0,"structures":[]}', "format 2 is newer than this boitata reads"),
    ('{"type":"Variogram","nugget":0.0,"structures":[]}', "missing or invalid format number"),
    ('[1, 2]', "expected a JSON object"),
    ('{"type":"Variogram","format":1,"nugget":0.0,"structures":[{"model":"Bogus","sill":1,"range":1}],'
     '"anisotropy":null}', "unknown variogram model"),
])
def test_envelope_errors(text, msg):
    with pytest.raises(InvalidInput, match=msg):
        Variogram.from_json(text)


VG = Variogram([("spherical", 1.0, 40.0)], nugget=0.1)
PASSES = [Search(30.0, max_samples=12, max_per_hole=4), Search(80.0, max_samples=8)]


def reference_file(path, name, method, n=50, fitted=True):
    """A file laid out as crates/py/src/persist.rs:114-133 writes one."""
    rng = np.random.default_rng(3)
    meta = {"type": name, "format": 1, "fitted": fitted, "method": method,
            "variogram": json.loads(VG.to_json()), "search": [json.loads(s.to_json()) for s in PASSES],
            "drift_data": [], "units": [None, None, None]}
    for d in (meta["variogram"], *meta["search"]):
        del d["type"], d["format"]
    hole = np.where(np.arange(n) % 5 == 0, np.nan, np.arange(n) // 7).astype(np.float64)
    cols = {"x": rng.uniform(0, 100, n), "y": rng.uniform(0, 100, n), "z": np.zeros(n),
            "value": rng.lognormal(0, 1, n), "hole": hole, "error_variance": np.zeros(n), "domain": np.full(n, np.nan)}
    if not fitted:
        cols = {k: v[:0] for k, v in cols.items()}
    arrays = [pa.array(v, mask=np.isnan(v) if k in ("hole", "domain") else None) for k, v in cols.items()]
    schema = pa.schema([pa.field(k, pa.float64()) for k in cols], metadata={"boitata": json.dumps(meta)})
    pq.write_table(pa.Table.from_arrays(arrays, schema=schema), path)
    return meta, cols


@pytest.mark.parametrize("cls,method,param", [
    (OrdinaryKriging, "Ordinary", None),
    (SimpleKriging, {"Simple": {"mean": 1.5}}, 1.5),
    (IndicatorKriging, {"Indicator": {"threshold": 0.8}}, 0.8),
])
def test_reads_a_reference_layout_file(tmp_path, cls, method, param):
    path = tmp_path / "model.parquet"
    meta, cols = reference_file(path, cls.__name__, {"Kriging": method})
    est = cls.from_parquet(path)
    assert type(est) is cls and len(est._passes) == 2
    if param is not None:
        assert est._param == param
    assert np.array_equal(est.values, cols["value"])
    for k in ("x", "y", "z", "error_variance"):
        assert np.array_equal(est._samples[k], cols[k])
    assert np.array_equal(est._samples["hole"], cols["hole"], equal_nan=True)
    # Written again: the same table and the same JSON.
    again = tmp_path / "again.parquet"
    est.to_parquet(again)
    a, b = pq.read_table(path), pq.read_table(again)
    assert a.column_names == b.column_names == ["x", "y", "z", "value", "hole", "error_variance", "domain"]
    for k in a.column_names:
        assert a[k].equals(b[k]), k
    assert json.loads(b.schema.metadata[b"boitata"]) == meta
    # Pickle carries the same state.
    back = pickle.loads(pickle.dumps(est))
    assert type(back) is cls and np.array_equal(back.values, est.values)
    assert np.array_equal(back._samples["hole"], est._samples["hole"], equal_nan=True)


def test_block_and_unfitted_round_trip(tmp_path):
    path = tmp_path / "block.parquet"
    bk = BlockKriging(VG, PASSES, (10, 10, 5), discretization=(4, 4, 2), precision="fp32_ir")
    bk.to_parquet(path)
    t = pq.read_table(path)
    assert t.num_rows == 0
    meta = json.loads(t.schema.metadata[b"boitata"])
    assert meta["fitted"] is False
    assert meta["method"] == {"Block": {"size": [10.0, 10.0, 5.0], "disc": {"nx": 4, "ny": 4, "nz": 2},
                                        "rotation": [0.0, 0.0, 0.0]}}
    back = BlockKriging.from_parquet(path, precision="fp32_ir")
    assert back._samples is None and bytes(back._block) == bytes(bk._block)
    with pytest.raises(InvalidInput, match="call fit first"):
        back.predict([[0.0, 0.0, 0.0]])
    p = pickle.loads(pickle.dumps(bk))
    assert p._precision == bk._precision and p._samples is None


def test_file_errors(tmp_path):
    path = tmp_path / "ok.parquet"
    reference_file(path, "OrdinaryKriging", {"Kriging": "Ordinary"})
    with pytest.raises(InvalidInput, match="expected a SimpleKriging"):
        SimpleKriging.from_parquet(path)
    with pytest.raises(FileError):
        OrdinaryKriging.from_parquet(tmp_path / "missing.parquet")
    # A method that is not the class's own.
    reference_file(path, "OrdinaryKriging", {"Kriging": {"Simple": {"mean": 0.0}}})
    with pytest.raises(InvalidInput, match="is not that of a OrdinaryKriging"):
        OrdinaryKriging.from_parquet(path)
    # A Parquet file without the metadata key.
    pq.write_table(pa.table({"x": [1.0]}), tmp_path / "plain.parquet")
    with pytest.raises(InvalidInput, match="no boitata metadata"):
        OrdinaryKriging.from_parquet(tmp_path / "plain.parquet")
    # Fitted columns with a null value.
    _, cols = reference_file(path, "OrdinaryKriging", {"Kriging": "Ordinary"})
    t = pq.read_table(path)
    v = cols["value"].copy()
    v[3] = np.nan
    t = t.set_column(3, pa.field("value", pa.float64()), pa.array(v, mask=np.isnan(v)))
    pq.write_table(t, path)
    with pytest.raises(InvalidInput, match="null in column value"):
        OrdinaryKriging.from_parquet(path)


def test_multiple_indicator_layout_and_reference_file(tmp_path):
    from boitata_amd.estimation import MultipleIndicatorKriging

    vg = Variogram([("spherical", 1.0, 30.0)])
    m = MultipleIndicatorKriging(vg, Search(25.0, max_samples=12), [0.5, 1.0], simple=True, tails=(0.0, 10.0),
                                 upper_tail=("power", 0.5))
    path = tmp_path / "mik.parquet"
    m.to_parquet(path)
    text = pq.read_table(path).schema.metadata[b"boitata"].decode()
    # crates/py/src/indicator.rs:115-125 with crates/estimation/src/indicator.rs:34-65.
    assert text == ('{"type":"MultipleIndicatorKriging","format":1,"fitted":false,"model":{"thresholds":[0.5,1.0],'
                    '"variograms":[{"nugget":0.0,"structures":[{"model":"Spherical","sill":1.0,"range":30.0}],'
                    '"anisotropy":null}],"simple":true,"tails":[0.0,10.0],"interpolation":"Global",'
                    '"upper_tail":{"Power":0.5}},"search":[{"min_samples":1,"max_samples":12,"radius":25.0,'
                    '"max_per_hole":null,"octant":false,"anisotropy":null,"high_grade":null}]}')
    # A fitted file with declustering weights, read and written again.
    meta = json.loads(text)
    meta["fitted"] = True
    n = 30
    rng = np.random.default_rng(5)
    cols = {"x": rng.uniform(0, 50, n), "y": rng.uniform(0, 50, n), "z": np.zeros(n), "value": rng.lognormal(0, 1, n),
            "hole": (np.arange(n) // 3).astype(np.float64), "error_variance": np.zeros(n),
            "domain": np.full(n, np.nan), "weight": rng.uniform(0.5, 2.0, n)}
    arrays = [pa.array(v, mask=np.isnan(v) if k == "domain" else None) for k, v in cols.items()]
    schema = pa.schema([pa.field(k, pa.float64()) for k in cols], metadata={"boitata": json.dumps(meta)})
    pq.write_table(pa.Table.from_arrays(arrays, schema=schema), path)
    back = MultipleIndicatorKriging.from_parquet(path)
    assert np.array_equal(back._weights, cols["weight"]) and np.array_equal(back._holes, np.arange(n) // 3)
    again = tmp_path / "again.parquet"
    back.to_parquet(again)
    a, b = pq.read_table(path), pq.read_table(again)
    assert a.column_names == b.column_names
    for k in a.column_names:
        assert a[k].equals(b[k]), k
    assert json.loads(b.schema.metadata[b"boitata"]) == meta
    p = pickle.loads(pickle.dumps(back))
    assert np.array_equal(p._xyz, back._xyz) and bytes(p._mik) == bytes(m._mik)


def test_with_search_keeps_the_samples(tmp_path):
    # python/boitata/estimation.py:357-375: a copy with other passes.
    path = tmp_path / "model.parquet"
    _, cols = reference_file(path, "SimpleKriging", {"Kriging": {"Simple": {"mean": 1.5}}})
    est = SimpleKriging.from_parquet(path)
    other = est.with_search(Search(55.0, max_samples=6))
    assert type(other) is SimpleKriging and other._param == 1.5 and len(other._passes) == 1
    assert other._samples is est._samples and other._engine is None and est.unit is None
    assert len(est._passes) == 2
    with pytest.raises(InvalidInput, match="calibrate it with target_efficiency"):
        est.with_search(Search(55.0, target_slope=0.9))
