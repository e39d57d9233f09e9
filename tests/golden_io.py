"""Loading the golden fixtures of tests/golden (written by
tests/golden/make_fixtures.py) and rebuilding their parameter structs."""
import os

import numpy as np

from boitata_amd import _abi

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with np.load(os.path.join(HERE, name), allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def search_case(p):
    ms, mx, radius, octant, mph, az, dip, rk, rs, rm = p.tolist()
    kw = {}
    if (az, dip, rk) != (0.0, 0.0, 0.0) or (rs, rm) != (1.0, 1.0):
        kw = dict(rotation=(az, dip, rk), ratios=(rs, rm))
    return _abi.make_search(int(ms), int(mx), radius, octant=bool(octant),
                            max_per_hole=None if mph < 0 else int(mph), **kw)


def c1_params(f):
    nug, sill, rng, az, dip, rk, rs, rm = f["variogram"].tolist()
    vg = _abi.make_variogram([("spherical", sill, rng)], nugget=nug, rotation=(az, dip, rk), ratios=(rs, rm))
    ms, mx, radius, saz, sdip, srk, srs, srm = f["search"].tolist()
    s = _abi.make_search(int(ms), int(mx), radius, rotation=(saz, sdip, srk), ratios=(srs, srm))
    return vg, s
