"""Host logic of the drop-in estimators (no device): parameter validation of
Variogram / Search as the reference's constructors
(crates/py/src/estimation.rs:415-531, crates/py/src/variogram.rs:329-362),
hole codes (crates/py/src/args.rs:278-292), shared locations
(args.rs:305-363), search passes (crates/estimation/src/batch.rs:132-147)."""
import math
import warnings

import numpy as np
import pytest

from boitata_amd import _abi
from boitata_amd.errors import InvalidInput
from boitata_amd.estimation import (CrossValidation, OrdinaryKriging, Search, SimpleKriging, Structure, Variogram,
                                    coords_of, distinct_rows, hole_codes, run_passes)


def test_search_from_ranges_sets_radius_and_ratios():
    s = Search(ranges=(400, 250, 80), rotation=(30, 20, 0), max_samples=32, min_samples=4)
    assert s._c.radius == 400.0
    assert s._c.has_anisotropy == 1
    assert math.isclose(s._c.ratio_semi, 250 / 400) and math.isclose(s._c.ratio_minor, 80 / 400)
    assert (s._c.azimuth, s._c.dip, s._c.rake) == (30.0, 20.0, 0.0)
    assert (s._c.min_samples, s._c.max_samples) == (4, 32)


def test_search_without_anisotropy_is_isotropic():
    s = Search(50.0)
    assert s._c.has_anisotropy == 0 and s._c.max_samples == 16 and s._c.min_samples == 1
    # Zero rotation and unit ratios are no anisotropy (variogram.rs:164-167).
    assert Search(50.0, rotation=(0, 0, 0), ratios=(1, 1))._c.has_anisotropy == 0


@pytest.mark.parametrize("kw,msg", [
    (dict(radius=10.0, ranges=(1, 1, 1)), "not both"),
    (dict(), "give radius or ranges"),
    (dict(ranges=(10, 5, 1), ratios=(0.5, 0.1)), "ratios with radius"),
    (dict(ranges=(10, 0, 1)), "ranges must be > 0"),
    (dict(radius=10.0, max_per_sector=2), "needs octant or sectors"),
    (dict(radius=10.0, max_samples=-1), "non-negative"),
])
def test_search_rejects(kw, msg):
    with pytest.raises(InvalidInput, match=msg):
        Search(**kw)


def test_search_checks_ranges_through_the_library():
    with pytest.raises(InvalidInput):
        Search(10.0, min_samples=8, max_samples=4)  # Search::check (search.rs:214-264)


def test_search_sector_caps():
    o = Search(10.0, octant=True, max_per_sector=3, max_samples=24)
    assert o._c.octant == 1 and o._c.max_per_octant == 3 and o._c.sectors == 0
    p = Search(10.0, sectors=6, max_samples=20)
    assert p._c.sectors == 6 and p._c.max_per_sector == 4  # ceil(20 / 6)
    q = Search(10.0, sectors=6, max_per_sector=2, max_samples=20)
    assert q._c.max_per_sector == 2 and q._c.max_per_octant == -1


@pytest.mark.parametrize("kw", [dict(soft=5.0), dict(high_grade=3.0), dict(length_unit="m")])
def test_search_features_out_of_scope(kw):
    with pytest.raises(NotImplementedError):
        Search(10.0, **kw)


def test_calibrated_search_arguments():
    # py/estimation.rs:458-472
    s = Search(10.0, target_slope=0.9)
    assert s._c.calibration == _abi.CALIBRATE_SLOPE and s._c.calibration_target == 0.9
    e = Search(10.0, target_efficiency=0.5)
    assert e._c.calibration == _abi.CALIBRATE_EFFICIENCY
    for kw in (dict(target_slope=0.9, target_efficiency=0.5), dict(target_slope=0.0), dict(target_slope=float("nan")),
               dict(target_efficiency=1.5)):
        with pytest.raises(InvalidInput):
            Search(10.0, **kw)


def test_variogram_tuples_and_structures():
    v = Variogram([("spherical", 0.8, 120), Structure("power", 1.0, 50.0, exponent=1.5)], nugget=0.2,
                  rotation=(30, 0, 0), ratios=(80 / 120, 30 / 120))
    c = v._c
    assert c.n_structures == 2 and c.nugget == 0.2 and c.has_anisotropy == 1
    assert c.structures[0].model == _abi.MODELS["spherical"] and c.structures[0].range == 120.0
    assert c.structures[1].model == _abi.MODELS["power"] and c.structures[1].param == 1.5
    assert not v.has_sill and math.isclose(v.sill, 2.0)


@pytest.mark.parametrize("args,kw,msg", [
    ([("spherical", 1.0, 10.0)], dict(nugget=-0.1), "nugget"),
    ([("spherical", 1.0)], {}, "tuples"),
    ([("nonsense", 1.0, 10.0)], {}, "unknown model"),
    ([("spherical", 1.0, 10.0)] * 5, {}, "at most"),
    ([("spherical", 1.0, -10.0)], {}, "variogram"),
])
def test_variogram_rejects(args, kw, msg):
    with pytest.raises(InvalidInput, match=msg):
        Variogram(args, **kw)


def test_simple_kriging_needs_a_sill():
    pw = Variogram([Structure("power", 1.0, 10.0, exponent=1.0)])
    with pytest.raises(InvalidInput, match="sill"):
        SimpleKriging(pw, Search(10.0), mean=1.0)


def test_estimators_need_a_variogram_and_searches():
    with pytest.raises(InvalidInput):
        OrdinaryKriging(None, Search(10.0))
    with pytest.raises(InvalidInput):
        OrdinaryKriging(Variogram([("spherical", 1.0, 10.0)]), [])
    with pytest.raises(InvalidInput):
        OrdinaryKriging(Variogram([("spherical", 1.0, 10.0)]), Search(10.0), precision="fp16")


def test_coords_two_columns_lie_on_zero_and_must_be_finite():
    a = coords_of([[1.0, 2.0], [3.0, 4.0]])
    assert a.shape == (2, 3) and (a[:, 2] == 0).all()
    with pytest.raises(InvalidInput, match="finite"):
        coords_of([[1.0, np.nan, 0.0]])
    with pytest.raises(InvalidInput, match="shape"):
        coords_of(np.zeros((3, 4)))


def test_hole_codes_follow_first_appearance():
    text, ids = hole_codes(["b", "a", None, "b", 7, "a"], 6)
    assert text == ["b", "a", None, "b", "7", "a"]
    assert ids.tolist() == [0, 1, 2, 0, 3, 1]
    with pytest.raises(InvalidInput):
        hole_codes(["a"], 2)


def test_shared_locations_keep_the_first():
    xyz = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [-0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [2.0, 0.0, 0.0]])
    keep, note = distinct_rows(xyz, ["h1", "h2", "h3", "h4", "h5"])
    assert keep.tolist() == [0, 1, 4]
    assert "holes h1, h3" in note and "holes h2, h4" in note
    keep, note = distinct_rows(xyz, None)
    assert "rows 0, 2" in note
    keep, note = distinct_rows(xyz[[0, 1, 4]], None)
    assert keep.tolist() == [0, 1, 2] and note is None


def test_passes_send_failures_onward():
    calls = []

    def one(s, rows):
        calls.append((s, rows.tolist()))
        ok = rows % 2 == 0 if s == "first" else np.ones(len(rows), dtype=bool)
        return rows * 10.0, rows * 1.0, ok

    v, q, which = run_passes(5, ["first", "second", "third"], one)
    assert calls == [("first", [0, 1, 2, 3, 4]), ("second", [1, 3])]
    assert v.tolist() == [0, 10, 20, 30, 40] and which.tolist() == [0, 1, 0, 1, 0]
    v, q, which = run_passes(2, ["x"], lambda s, r: (r * 1.0, r * 1.0, np.zeros(len(r), dtype=bool)))
    assert np.isnan(v).all() and (which == -1).all()


def test_cross_validation_statistics():
    cv = CrossValidation(np.array([1.0, 2.0, 3.0, 4.0]), np.array([1.5, 2.0, np.nan, 3.5]),
                         np.array([0.25, 1.0, np.nan, 0.25]))
    assert math.isclose(cv.mean_error, 0.0)
    assert math.isclose(cv.rmse, math.sqrt((0.25 + 0 + 0.25) / 3))
    assert math.isclose(cv.standardized_squared_error, (1 + 0 + 1) / 3)
    assert math.isclose(cv.slope, np.polyfit([1.5, 2.0, 3.5], [1.0, 2.0, 4.0], 1)[0])


def test_fit_warns_on_shared_locations_before_any_device_call(monkeypatch):
    est = OrdinaryKriging(Variogram([("spherical", 1.0, 10.0)]), Search(10.0))

    class Fake:
        def set_samples(self, x, y, z, v, err, hole):
            self.got = (x, y, z, v, err, hole)

    fake = Fake()
    monkeypatch.setattr(est, "_device_engine", lambda: fake)
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        est.fit([[0, 0, 0], [0, 0, 0], [1, 1, 1]], [1.0, 2.0, 3.0], holes=["a", "b", "c"])
    assert any("holes a, b" in str(x.message) for x in w)
    assert fake.got[3].tolist() == [1.0, 3.0] and fake.got[5].tolist() == [0, 2]
    assert est.values.tolist() == [1.0, 3.0]
    with pytest.raises(NotImplementedError):
        est.predict([[0, 0, 0]], anisotropy=object())


def test_simple_kriging_takes_no_slope_calibration():
    vg = Variogram([("spherical", 1.0, 50.0)], nugget=0.1)
    with pytest.raises(InvalidInput):
        SimpleKriging(vg, Search(30.0, target_slope=0.9), mean=1.0)
    SimpleKriging(vg, Search(30.0, target_efficiency=0.5), mean=1.0)


def test_multiple_indicator_kriging_checks_its_parameters():
    # crates/py/src/indicator.rs:131-192 and indicator.rs:291-317.
    from boitata_amd.estimation import MultipleIndicatorKriging

    vg = Variogram([("spherical", 1.0, 30.0)])
    s = Search(20.0)
    with pytest.raises(InvalidInput, match="strictly increasing"):
        MultipleIndicatorKriging(vg, s, [1.0, 1.0])
    with pytest.raises(InvalidInput, match="one per threshold"):
        MultipleIndicatorKriging([vg, vg], s, [0.0, 1.0, 2.0])
    with pytest.raises(InvalidInput, match="finite threshold"):
        MultipleIndicatorKriging(vg, s, [])
    with pytest.raises(InvalidInput, match="interpolation"):
        MultipleIndicatorKriging(vg, s, [1.0], interpolation="cubic")
    with pytest.raises(InvalidInput, match="upper tail"):
        MultipleIndicatorKriging(vg, s, [1.0], upper_tail=("lognormal", 1.0))
    with pytest.raises(InvalidInput, match="exponent"):
        MultipleIndicatorKriging(vg, s, [1.0], upper_tail=("power", 0.0))
    with pytest.raises(InvalidInput, match="last threshold > 0"):
        MultipleIndicatorKriging(vg, s, [-1.0], upper_tail=("hyperbolic", 2.0))
    with pytest.raises(InvalidInput, match="tails"):
        MultipleIndicatorKriging(vg, s, [1.0], tails=(2.0, 1.0))
    with pytest.raises(InvalidInput, match="target_slope"):
        MultipleIndicatorKriging(vg, Search(20.0, target_efficiency=0.5), [1.0])
    m = MultipleIndicatorKriging([vg, vg], [s, Search(40.0)], [0.5, 1.5], upper_tail=("power", 2.0))
    assert m.thresholds == [0.5, 1.5]
    with pytest.raises(InvalidInput, match="fit"):
        m.predict([[0.0, 0.0]])
