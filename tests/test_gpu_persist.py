"""Reloaded estimators predict bit for bit as the saved ones
(python/tests/test_persist.py:209-224 of the reference): to_parquet /
from_parquet and pickle of fitted and unfitted estimators on the 5,000-sample
drill-hole workload, with holes (some samples untagged), measurement error
variances, search passes, every kriging kind, block support and both
precisions; estimates, variances, the diagnostics table and leave-one-out."""
import pickle

import numpy as np
import pytest

from boitata_amd import datasets
from boitata_amd.estimation import (BlockKriging, IndicatorKriging, OrdinaryKriging, Search, SimpleKriging,
                                    Variogram)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c1():
    w = datasets.config1()
    P = np.stack([w.x, w.y, w.z], 1)
    holes = [None if i % 9 == 0 else f"DH{h}" for i, h in enumerate(w.hole)]
    err = np.where(np.arange(len(w.x)) % 4 == 0, 0.05, 0.0)
    return w, P, holes, err, w.targets[::37].copy()


def same(a, b):
    for x, y in zip(a, b):
        x, y = np.asarray(x), np.asarray(y)
        assert x.shape == y.shape and np.array_equal(x, y, equal_nan=True)


VG = Variogram([("spherical", 0.8, 120.0), ("exponential", 0.3, 60.0)], nugget=0.1, rotation=(30, 0, 0),
               ratios=(80 / 120, 30 / 120))
PASSES = [Search(ranges=(120, 80, 30), rotation=(30, 0, 0), min_samples=6, max_samples=24, max_per_hole=6,
                 octant=True, max_per_sector=4), Search(300.0, max_samples=12, min_samples=1)]


def estimators(precision):
    return [
        (OrdinaryKriging(VG, PASSES, precision=precision), True),
        (SimpleKriging(VG, PASSES, mean=1.0, precision=precision), True),
        (IndicatorKriging(VG, PASSES, 1.2, precision=precision), True),
        (BlockKriging(VG, PASSES, (10, 10, 5), discretization=(3, 3, 2), precision=precision), False),
    ]


def outputs(est, T, cv):
    v, q = est.predict(T, return_variance=True)
    d = est.predict(T, diagnostics=True)
    out = [v, q] + [d[c] for c in d.column_names]
    if cv:
        c = est.cross_validate()
        out += [c.estimate, c.variance]
    return out


@pytest.mark.parametrize("precision", ["fp64", "fp32_ir"])
@pytest.mark.parametrize("k", range(4))
def test_round_trip_predicts_bit_identically(c1, tmp_path, precision, k):
    w, P, holes, err, T = c1
    est, cv = estimators(precision)[k]
    kind = type(est).__name__
    print(f"[persist] {kind} {precision}", flush=True)
    path = tmp_path / "model.parquet"
    est.to_parquet(path)
    unfitted = type(est).from_parquet(path, precision=precision)
    extra = dict(holes=holes) if kind == "BlockKriging" else dict(holes=holes, error_variance=err)
    est.fit(P, w.value, **extra)
    ref = outputs(est, T, cv)
    assert np.isfinite(ref[0]).mean() > 0.9
    same(outputs(unfitted.fit(P, w.value, **extra), T, cv), ref)
    est.to_parquet(path)
    for back in (type(est).from_parquet(path, precision=precision), pickle.loads(pickle.dumps(est))):
        assert type(back) is type(est)
        same(outputs(back, T, cv), ref)


@pytest.mark.parametrize("precision", ["fp64", "fp32_ir"])
def test_multiple_indicator_round_trip(c1, tmp_path, precision):
    # python/tests/test_persist.py:180-190: the conditional distributions, the
    # diagnostics table and leave-one-out with the PIT, after a reload.
    from boitata_amd.estimation import MultipleIndicatorKriging

    w, P, holes, _, T = c1
    print(f"[persist] MultipleIndicatorKriging {precision}", flush=True)
    thr = list(np.quantile(w.value, [0.2, 0.4, 0.6, 0.8]))
    vgs = [Variogram([("spherical", 0.2, 100.0 + 20 * k)], nugget=0.05) for k in range(4)]
    weights = 1.0 + (np.arange(len(w.x)) % 3)
    est = MultipleIndicatorKriging(vgs, PASSES, thr, tails=(0.0, float(w.value.max()) * 1.5),
                                   upper_tail=("hyperbolic", 1.5), precision=precision)

    def run(e):
        s = e.predict(T, cutoffs=[thr[1], thr[3]], quantiles=[0.1, 0.5, 0.9])
        d = e.predict(T, cutoffs=[thr[1]], diagnostics=True).diagnostics
        c = e.cross_validate()
        return ([s.cdf, s.correction, s.mean, s.variance, s.probability_above, s.mean_above, s.quantile_values]
                + [d[k] for k in d.column_names] + [c.estimate, c.variance, c.cdf, c.pit])

    path = tmp_path / "mik.parquet"
    est.to_parquet(path)
    unfitted = MultipleIndicatorKriging.from_parquet(path, precision=precision)
    est.fit(P, w.value, weights=weights, holes=holes)
    ref = run(est)
    assert np.isfinite(ref[2]).mean() > 0.9
    same(run(unfitted.fit(P, w.value, weights=weights, holes=holes)), ref)
    est.to_parquet(path)
    for back in (MultipleIndicatorKriging.from_parquet(path, precision=precision), pickle.loads(pickle.dumps(est))):
        same(run(back), ref)


def test_with_search_predicts_as_a_fresh_estimator(c1):
    w, P, holes, err, T = c1
    est = OrdinaryKriging(VG, PASSES).fit(P, w.value, holes=holes, error_variance=err)
    s = Search(200.0, max_samples=10, min_samples=2)
    fresh = OrdinaryKriging(VG, s).fit(P, w.value, holes=holes, error_variance=err)
    same(est.with_search(s).predict(T, return_variance=True), fresh.predict(T, return_variance=True))
