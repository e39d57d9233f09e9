"""The no-allocation entry point (bk_krige_noalloc) against bk_krige through
the C ABI: the same kernels on views of the library's static device array,
so statuses and neighbour counts are equal and values and variances equal to
the last bit. Two calls in a row, the second with a smaller input; the limits
are refused before any device work."""
import numpy as np
import pytest

from boitata_amd import _abi, datasets

pytestmark = pytest.mark.gpu


def dev(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).to(torch.device("cuda:0"))


def host(out):
    return {k: v.cpu().numpy() for k, v in out.items()}


def same(a, b):
    for k in ("status", "n_used"):
        assert (a[k] == b[k]).all(), k
    ok = a["status"] == 0
    for k in ("value", "variance", "lagrange"):
        assert (a[k][ok].view(np.uint64) == b[k][ok].view(np.uint64)).all(), k
        assert np.isnan(a[k][~ok]).all() and np.isnan(b[k][~ok]).all(), k


@pytest.mark.parametrize("precision", [_abi.FP64, _abi.FP32_IR])
def test_noalloc_equals_krige_twice(engine, precision):
    w = datasets.config2()  # 100k samples, <= 32 neighbours
    engine.set_samples(dev(w.x), dev(w.y), dev(w.z), dev(w.value), None, dev(w.hole.astype(np.int32)))
    for m in (_abi.STATIC_MAX_TARGETS, 5000):  # the largest call, then a smaller one
        T = w.targets[:m]
        tx, ty, tz = dev(T[:, 0]), dev(T[:, 1]), dev(T[:, 2])
        a = host(engine.krige_noalloc(w.variogram, w.search, tx, ty, tz, precision=precision))
        b = host(engine.krige(w.variogram, w.search, tx, ty, tz, precision=precision))
        assert (a["status"] == 0).mean() > 0.9
        same(a, b)


def test_noalloc_capped_search_and_simple_kriging(engine):
    w = datasets.config3(target_counts=(40, 40, 20))
    engine.set_samples(dev(w.x), dev(w.y), dev(w.z), dev(w.value))
    T = w.targets[::3]
    tx, ty, tz = dev(T[:, 0]), dev(T[:, 1]), dev(T[:, 2])
    for kind, param in ((_abi.ORDINARY, 0.0), (_abi.SIMPLE, 1.0)):
        a = host(engine.krige_noalloc(w.variogram, w.search, tx, ty, tz, kind, param, _abi.FP32_IR))
        b = host(engine.krige(w.variogram, w.search, tx, ty, tz, kind, param, _abi.FP32_IR))
        same(a, b)


def test_noalloc_limits(engine):
    from boitata_amd.engine import EngineError

    w = datasets.config1()
    engine.set_samples(dev(w.x), dev(w.y), dev(w.z), dev(w.value))
    big = dev(np.zeros(_abi.STATIC_MAX_TARGETS + 1))
    with pytest.raises(EngineError):
        engine.krige_noalloc(w.variogram, w.search, big, big, big)
    wide = _abi.make_search(1, 65, 100.0)
    small = dev(np.zeros(4))
    with pytest.raises(EngineError):
        engine.krige_noalloc(w.variogram, wide, small, small, small)
    with pytest.raises(EngineError):  # host arrays would need staging
        engine.krige_noalloc(w.variogram, w.search, np.zeros(4), np.zeros(4), np.zeros(4))
