"""Multiple indicator kriging and the diagnostics table at the full size of
the 1M-target workload (100k drill-hole composites, <= 32 neighbours): device
arrays throughout, the oracle on an evenly strided subset, and properties that
hold at any size over the whole pass:
- the FP64 and FP32+IR passes give the same statuses and agree to 1e-6;
- corrected distributions lie in [0, 1] and do not decrease; quantiles do
  not decrease; probabilities above the cutoffs do not increase;
- with one variogram the first threshold's kriged probability is the
  indicator kriging estimate (before its clamp), to 1e-12;
- diagnostics: holes never outnumber samples, the table's estimates are the
  pass's estimates, the mean distance stays inside the search ellipsoid's
  major range."""
import numpy as np
import pytest

import oracle
from boitata_amd import _abi, datasets

pytestmark = pytest.mark.gpu


def dev(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).to(torch.device("cuda:0"))


def host(out):
    return {k: (v.cpu().numpy() if hasattr(v, "cpu") else v) for k, v in out.items()}


@pytest.fixture(scope="module")
def c2():
    return datasets.config2()


def test_mik_full(engine, c2):
    w = c2
    engine.set_samples(dev(w.x), dev(w.y), dev(w.z), dev(w.value), None, dev(w.hole.astype(np.int32)))
    thr = np.quantile(w.value, np.linspace(0.1, 0.9, 9))
    p = _abi.make_mik(thr, 1)
    kw = dict(cutoffs=[0.5, 1.0, 2.0], quantiles=[0.1, 0.5, 0.9])
    T = w.targets
    tx, ty, tz = dev(T[:, 0]), dev(T[:, 1]), dev(T[:, 2])
    g64 = host(engine.mik_predict(p, thr, [w.variogram], w.search, tx, ty, tz, precision=_abi.FP64, **kw))
    g32 = host(engine.mik_predict(p, thr, [w.variogram], w.search, tx, ty, tz, precision=_abi.FP32_IR, **kw))
    assert np.array_equal(g64["status"], g32["status"])
    ok = g64["status"] == 0
    assert ok.mean() > 0.99
    for k in ("cdf", "mean", "variance", "probability_above"):
        a, b = g32[k][..., ok], g64[k][..., ok]
        assert (np.abs(a - b) / np.maximum(np.abs(b), 1e-3)).max() <= 1e-6, k
    cdf = g64["cdf"][:, ok]
    assert np.all((cdf >= 0) & (cdf <= 1)) and np.all(np.diff(cdf, axis=0) >= 0)
    assert np.all(np.diff(g64["quantile_values"][:, ok], axis=0) >= 0)
    assert np.all(np.diff(g64["probability_above"][:, ok], axis=0) <= 0)
    assert np.isnan(g64["mean"][~ok]).all() and not np.isnan(g64["mean"][ok]).any()
    # One variogram: threshold 0's kriged probability is indicator kriging's sum.
    ik = host(engine.krige(w.variogram, w.search, tx, ty, tz, kind=_abi.INDICATOR, param=float(thr[0])))
    raw0 = g64["raw"][0, ok]
    assert np.abs(np.clip(raw0, 0, 1) - ik["value"][ok]).max() <= 1e-12
    # The oracle on every 997th target.
    sub = np.arange(0, len(T), 997)
    d = oracle.Data(w.x, w.y, w.z, w.value, None, w.hole)
    o = oracle.mik_predict(d, p, thr, [w.variogram], w.search, T[sub], **kw)
    assert np.array_equal(o["status"], g64["status"][sub])
    oks = o["status"] == 0
    for k in ("raw", "cdf", "mean", "variance", "probability_above"):
        a, b = g64[k][..., sub][..., oks], o[k][..., oks]
        assert (np.abs(a - b) / np.maximum(np.abs(b), 1e-3)).max() <= 1e-10, k


def test_diagnostics_full(engine, c2):
    w = c2
    engine.set_samples(dev(w.x), dev(w.y), dev(w.z), dev(w.value), None, dev(w.hole.astype(np.int32)))
    T = w.targets
    tx, ty, tz = dev(T[:, 0]), dev(T[:, 1]), dev(T[:, 2])
    g = host(engine.krige_diagnostics(w.variogram, w.search, tx, ty, tz, precision=_abi.FP32_IR))
    k = host(engine.krige(w.variogram, w.search, tx, ty, tz, precision=_abi.FP32_IR))
    assert np.array_equal(g["status"], k["status"])
    ok = g["status"] == 0
    assert np.array_equal(g["value"][ok], k["value"][ok]) and np.array_equal(g["n_used"], k["n_used"])
    assert np.all(g["n_holes"][ok] >= 1) and np.all(g["n_holes"][ok] <= g["n_used"][ok])
    assert np.all(g["mean_distance"][ok] > 0) and np.all(g["mean_distance"][ok] <= w.search.radius)
    assert np.all(g["negative_weight_sum"][ok] <= 0) and np.all(g["target_met"] == -1)
    sub = np.arange(0, len(T), 1999)
    d = oracle.Data(w.x, w.y, w.z, w.value, None, w.hole)
    o = oracle.krige_diag(d, w.variogram, w.search, T[sub])
    assert np.array_equal(o["status"], g["status"][sub])
    oks = o["status"] == 0
    for key in ("n_used", "n_holes"):
        assert np.array_equal(o[key][oks], g[key][sub][oks]), key
    assert np.abs(o["mean_distance"][oks] - g["mean_distance"][sub][oks]).max() <= 1e-13 * w.search.radius
