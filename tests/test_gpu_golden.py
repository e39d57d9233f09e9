"""The HIP path against the committed golden fixtures (tests/golden), through
the C ABI: neighbour sets bit for bit, FP64 estimates at 1e-10, FP32 +
refinement at 1e-6 (relative, floor 1e-3 of the sill)."""
import numpy as np
import pytest

from boitata_amd import _abi
from golden_io import c1_params, load, search_case

pytestmark = pytest.mark.gpu


def test_lcg_search_fixture(engine):
    f = load("lcg_search.npz")
    engine.set_samples(f["x"], f["y"], f["z"], f["value"], None, f["hole"])
    T = f["targets"]
    for c in range(4):
        s = search_case(f[f"case{c}_params"])
        idx, cnt, st = engine.neighbors(s, None, T[:, 0], T[:, 1], T[:, 2])
        assert (cnt == f[f"case{c}_count"]).all() and (st == f[f"case{c}_status"]).all()
        for t in range(len(cnt)):
            assert (idx[t, :cnt[t]] == f[f"case{c}_idx"][t, :cnt[t]]).all(), (c, t)


@pytest.mark.parametrize("precision,tol", [(_abi.FP64, 1e-10), (_abi.FP32_IR, 1e-6)])
def test_c1_kriging_fixture(engine, precision, tol):
    f = load("c1_kriging.npz")
    vg, s = c1_params(f)
    engine.set_samples(f["x"], f["y"], f["z"], f["value"], None, f["hole"])
    T = f["targets"]
    floor = 1e-3 * (vg.nugget + vg.structures[0].sill)
    g = engine.krige(vg, s, T[:, 0], T[:, 1], T[:, 2], precision=precision)
    assert (g["status"] == f["status"]).all()
    ok = f["status"] == 0
    for k, ref in (("value", "est"), ("variance", "variance")):
        err = np.abs(g[k][ok] - f[ref][ok]) / np.maximum(np.abs(f[ref][ok]), floor)
        assert err.max(initial=0.0) <= tol, (k, err.max())
    gb = engine.block_krige(vg, s, _abi.make_block((10.0, 10.0, 10.0), (2, 2, 2)), T[:, 0], T[:, 1], T[:, 2],
                            precision=precision)
    assert (gb["status"] == f["block_status"]).all()
    okb = f["block_status"] == 0
    err = np.abs(gb["value"][okb] - f["block_est"][okb]) / np.maximum(np.abs(f["block_est"][okb]), floor)
    assert err.max(initial=0.0) <= tol
