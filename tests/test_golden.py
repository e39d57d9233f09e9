"""The oracle against the committed golden fixtures (tests/golden): seeded
inputs (the reference's LCG clouds, a C1 drill-hole subset) whose outputs were
cross-checked at write time against brute-force selection and dense numpy
solves. A change of the oracle's results shows here."""
import numpy as np

import oracle
from boitata_amd import _abi
from golden_io import c1_params, load, search_case


def test_lcg_search_fixture():
    f = load("lcg_search.npz")
    d = oracle.Data(f["x"], f["y"], f["z"], f["value"], None, f["hole"])
    for c in range(4):
        s = search_case(f[f"case{c}_params"])
        idx, cnt, st = oracle.neighbors(d, s, None, f["targets"])
        assert (cnt == f[f"case{c}_count"]).all() and (st == f[f"case{c}_status"]).all()
        for t in range(len(cnt)):
            assert (idx[t, :cnt[t]] == f[f"case{c}_idx"][t, :cnt[t]]).all(), (c, t)


def test_c1_kriging_fixture():
    f = load("c1_kriging.npz")
    vg, s = c1_params(f)
    d = oracle.Data(f["x"], f["y"], f["z"], f["value"], None, f["hole"])
    o = oracle.krige(d, vg, s, f["targets"])
    assert (o["status"] == f["status"]).all()
    ok = o["status"] == 0
    for k, ref in (("value", "est"), ("variance", "variance"), ("lagrange", "lagrange")):
        assert np.array_equal(o[k][ok], f[ref][ok]), k
    ob = oracle.block_krige(d, vg, s, _abi.make_block((10.0, 10.0, 10.0), (2, 2, 2)), f["targets"])
    okb = ob["status"] == 0
    assert (ob["status"] == f["block_status"]).all()
    assert np.array_equal(ob["value"][okb], f["block_est"][okb])
