"""Sequential Gaussian simulation of normal scores on the MI355X engine.

``SGS(variogram, search)`` keeps the constructor of the reference's ``SGS``
(crates/py/src/simulation.rs:1124-1260) for one search pass; ``fit`` takes
the conditioning data already in normal scores and ``simulate`` returns score
realizations, one row per realization. Each realization is the node loop of
crates/simulation/src/sgs.rs:600-750 (search the data and the nodes simulated
before, simple-krige with mean 0, draw, add the node), computed by
``bk_sgs``: a node on a datum takes the datum's score, a node without
neighbours draws from the marginal, the variogram is scaled to a unit sill.

Not covered: the normal-score transform and its back-transform (declustering
weights, trend classes, domains), summaries, local anisotropy, secondary
variables and several search passes. The visiting order and the standard
normal draws come from numpy's generator (``numpy.random.default_rng(seed)``,
a permutation then normal draws), not from the reference's generator; a
caller may pass its own ``path`` and ``draws``.
"""
from __future__ import annotations

import numpy as np

from . import _abi
from .errors import InvalidInput
from .estimation import _PRECISIONS, Search, Variogram


def _coords(a, what):
    a = np.asarray(a, dtype=np.float64)
    if a.ndim != 2 or a.shape[1] not in (2, 3):
        raise InvalidInput(f"{what}: expected an (n, 2) or (n, 3) array")
    if a.shape[1] == 2:
        a = np.concatenate([a, np.zeros((len(a), 1))], 1)
    if not np.isfinite(a).all():
        raise InvalidInput(f"{what} must be finite")
    return np.ascontiguousarray(a)


class SGS:
    """Sequential Gaussian simulation of normal scores (see the module)."""

    def __init__(self, variogram: Variogram, search: Search, *, precision: str = "fp64", device: int = 0):
        if isinstance(search, (list, tuple)):
            if len(search) != 1:
                raise NotImplementedError("one search pass is supported by this engine")
            search = search[0]
        if precision not in _PRECISIONS:
            raise InvalidInput(f"precision is one of {', '.join(_PRECISIONS)}")
        self._variogram = variogram
        self._search = search
        self._precision = _PRECISIONS[precision]
        self._device = device
        self._engine = None
        self._n = None

    @property
    def variogram(self) -> Variogram:
        return self._variogram

    @property
    def search(self) -> list[Search]:
        return [self._search]

    @property
    def n_samples(self):
        """Number of fitted samples; None before `fit`."""
        return self._n

    def fit(self, coords, scores, holes=None) -> "SGS":
        """Conditioning data: locations (n, 2 | 3), their normal scores, and
        optional integer hole codes (for max_per_hole)."""
        xyz = _coords(coords, "coords")
        s = np.ascontiguousarray(scores, dtype=np.float64)
        if len(s) != len(xyz):
            raise InvalidInput("one score per sample")
        if len(s) == 0:
            raise InvalidInput("no conditioning data")
        if not np.isfinite(s).all():
            raise InvalidInput("scores must be finite")
        h = None if holes is None else np.ascontiguousarray(holes, dtype=np.int32)
        if self._engine is None:
            from .engine import Engine

            self._engine = Engine(self._device)
        self._engine.set_samples(xyz[:, 0], xyz[:, 1], xyz[:, 2], s, None, h)
        self._n = len(s)
        return self

    def simulate(self, targets, n: int = 1, seed: int = 0, path=None, draws=None) -> np.ndarray:
        """`n` realizations at the targets (m, 2 | 3), as an (n, m) array of
        scores. Realization r visits the nodes in `path[r]` and draws
        `draws[r]` (the k-th node of the path not on a datum takes
        draws[r][k]), or, without them, a permutation and m normal draws from
        numpy.random.default_rng(seed + r)."""
        if self._engine is None:
            raise InvalidInput("call fit first")
        g = _coords(targets, "targets")
        m = len(g)
        out = np.empty((n, m))
        self.sources = np.empty((n, m), dtype=np.int32)
        for r in range(n):
            if path is None or draws is None:
                rng = np.random.default_rng(seed + r)
                p = rng.permutation(m)
                z = rng.normal(size=m)
            else:
                p = np.asarray(path[r], dtype=np.int64)
                z = np.asarray(draws[r], dtype=np.float64)
            res = self._engine.sgs(self._variogram._c, self._search._c, g[:, 0], g[:, 1], g[:, 2], p, z,
                                   precision=self._precision)
            if (res["source"] == -3).any():
                raise InvalidInput("a node's kriging system failed (singular)")
            out[r] = res["score"]
            self.sources[r] = res["source"]
        return out

    def close(self):
        if self._engine is not None:
            self._engine.close()
            self._engine = None
