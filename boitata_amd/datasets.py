"""Seeded synthetic workloads of the kriging hot path (drill-hole composites and
block-model targets).

Values are exp(mu + sigma Z) with Z a standard Gaussian field drawn by random
Fourier features from the spectral density of the workload's variogram
(spherical approximated by a Gaussian of the same practical range), so the
data have the intended anisotropic continuity; sigma^2 = ln(1 + CV^2) and
mu = -sigma^2 / 2 give a lognormal mean of 1. Every engine sees the same
arrays, so the field method does not affect any comparison.

Targets are block centroids in x-fastest order (crates/core/src/block_model.rs:62-86).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import _abi


@dataclass
class Workload:
    name: str
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    value: np.ndarray
    hole: np.ndarray | None
    targets: np.ndarray | None           # (m, 3)
    variogram: _abi.Variogram
    variogram_shared: _abi.Variogram     # the reference-expressible form (shared ratios)
    search: _abi.Search
    block: _abi.Block | None = None
    notes: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return len(self.x)


def rotation_matrix(az, dip, rake):
    """crates/core/src/rotation.rs:8-16 (for data generation only)."""
    a, d, r = np.radians([az, dip, rake])
    sa, ca, sd, cd, sr, cr = np.sin(a), np.cos(a), np.sin(d), np.cos(d), np.sin(r), np.cos(r)
    r1 = np.array([[sa, ca, 0], [-ca, sa, 0], [0, 0, 1.0]])
    r2 = np.array([[cd, 0, -sd], [0, 1.0, 0], [sd, 0, cd]])
    r3 = np.array([[1.0, 0, 0], [0, cr, sr], [0, -sr, cr]])
    return r3 @ r2 @ r1


def gaussian_field(pts, structures, nugget, rotation, rng, features=384):
    """Standard Gaussian field at pts (n, 3) with the nested model's continuity.

    structures: [(model, sill, (major, semi, minor))].
    """
    R = rotation_matrix(*rotation)
    total = nugget + sum(s[1] for s in structures)
    out = np.zeros(len(pts))
    for model, sill, ranges in structures:
        k = max(int(features * sill / max(total - nugget, 1e-12)), 16)
        scale = np.diag(1.0 / np.asarray(ranges, dtype=np.float64)) @ R
        u = pts @ scale.T  # lags in units of the practical range
        if model == "exponential":
            # exp(-3h): spectral density of a multivariate Cauchy (t, 1 dof) scaled by 3.
            g = rng.standard_normal((k, 3))
            w = 3.0 * g / np.abs(rng.standard_normal((k, 1)))
        else:
            # Gaussian exp(-3h^2) (and spherical, approximated): normal with sd sqrt(6).
            w = np.sqrt(6.0) * rng.standard_normal((k, 3))
        phase = rng.uniform(0, 2 * np.pi, k)
        block = 16384
        acc = np.zeros(len(pts))
        for s in range(0, len(pts), block):
            acc[s:s + block] = np.cos(u[s:s + block] @ w.T + phase).sum(axis=1)
        out += np.sqrt(sill / total) * np.sqrt(2.0 / k) * acc
    out += np.sqrt(nugget / total) * rng.standard_normal(len(pts))
    return out


def lognormal(z, cv=1.5):
    s2 = np.log(1.0 + cv * cv)
    return np.exp(-s2 / 2.0 + np.sqrt(s2) * (z - z.mean()) / z.std())


def drillholes(nx, ny, spacing, jitter, depth, step, rng, origin=(0.0, 0.0)):
    """Vertical holes on an nx x ny grid, jittered; composites every `step` m to -depth."""
    cx = origin[0] + spacing[0] * (np.arange(nx) + 0.5)
    cy = origin[1] + spacing[1] * (np.arange(ny) + 0.5)
    X, Y = np.meshgrid(cx, cy, indexing="xy")
    X = X.ravel() + rng.uniform(-jitter, jitter, X.size)
    Y = Y.ravel() + rng.uniform(-jitter, jitter, Y.size)
    zc = -(np.arange(int(round(depth / step))) * step + step / 2.0)
    x = np.repeat(X, len(zc))
    y = np.repeat(Y, len(zc))
    z = np.tile(zc, len(X))
    hole = np.repeat(np.arange(len(X), dtype=np.int32), len(zc))
    return x, y, z, hole


def block_centroids(counts, size, origin=(0.0, 0.0, 0.0)):
    """Centroids of a block model, x fastest (block_model.rs:62-86, unrotated)."""
    nx, ny, nz = counts
    i = np.arange(nx * ny * nz)
    ix, iy, iz = i % nx, (i // nx) % ny, i // (nx * ny)
    return np.stack([origin[0] + (ix + 0.5) * size[0], origin[1] + (iy + 0.5) * size[1],
                     origin[2] + (iz + 0.5) * size[2]], axis=1)


def config1(seed=11, targets=True) -> Workload:
    """C1: 50 holes (10 x 5 over 500 x 500 m, +-25 m), 2 m composites to -200 m
    (5,000 samples); spherical nugget 0.2 + sill 0.8, ranges 120/80/30, az 30;
    50 x 50 x 20 targets at 10 m; ellipsoid search, 4..24."""
    rng = np.random.default_rng(seed)
    x, y, z, hole = drillholes(10, 5, (50.0, 100.0), 25.0, 200.0, 2.0, rng)
    pts = np.stack([x, y, z], 1)
    zf = gaussian_field(pts - pts.mean(0), [("spherical", 0.8, (120.0, 80.0, 30.0))], 0.2, (30.0, 0.0, 0.0), rng)
    value = lognormal(zf)
    vg = _abi.make_variogram([("spherical", 0.8, 120.0)], nugget=0.2, rotation=(30.0, 0.0, 0.0),
                             ratios=(80.0 / 120.0, 30.0 / 120.0))
    search = _abi.make_search(min_samples=4, max_samples=24, radius=120.0, rotation=(30.0, 0.0, 0.0),
                              ratios=(80.0 / 120.0, 30.0 / 120.0))
    t = block_centroids((50, 50, 20), (10.0, 10.0, 10.0), origin=(0.0, 0.0, -200.0)) if targets else None
    return Workload("C1 ordinary point kriging, 5k samples, 50k targets, <=24 neighbours", x, y, z, value, hole, t,
                    vg, vg, search)


def _c2_variograms():
    per = _abi.make_variogram([("exponential", 0.5, 150.0, None, (100.0 / 150.0, 40.0 / 150.0)),
                               ("spherical", 0.4, 400.0, None, (250.0 / 400.0, 80.0 / 400.0))],
                              nugget=0.1, rotation=(30.0, 20.0, 0.0), ratios=(0.625, 0.2))
    shared = _abi.make_variogram([("exponential", 0.5, 150.0), ("spherical", 0.4, 400.0)], nugget=0.1,
                                 rotation=(30.0, 20.0, 0.0), ratios=(0.625, 0.2))
    return per, shared


def config2(seed=12, holes=(20, 20), targets=True, target_counts=(200, 200, 25)) -> Workload:
    """C2: 400 holes (20 x 20 at 100 m, +-50 m) over 2000 x 2000 m, 250 composites
    of 2 m (100,000 samples); nugget 0.1 + exponential (0.5; 150/100/40) +
    spherical (0.4; 400/250/80), rotation (30, 20, 0); 200 x 200 x 25 blocks of
    10 x 10 x 20 m (1,000,000 targets); ellipsoid search 400/250/80, 4..32."""
    rng = np.random.default_rng(seed)
    sx = 2000.0 / holes[0]
    sy = 2000.0 / holes[1]
    x, y, z, hole = drillholes(holes[0], holes[1], (sx, sy), min(sx, sy) / 2.0, 500.0, 2.0, rng)
    pts = np.stack([x, y, z], 1)
    zf = gaussian_field(pts - pts.mean(0), [("exponential", 0.5, (150.0, 100.0, 40.0)),
                                            ("spherical", 0.4, (400.0, 250.0, 80.0))], 0.1, (30.0, 20.0, 0.0), rng)
    value = lognormal(zf)
    per, shared = _c2_variograms()
    search = _abi.make_search(min_samples=4, max_samples=32, radius=400.0, rotation=(30.0, 20.0, 0.0),
                              ratios=(250.0 / 400.0, 80.0 / 400.0))
    t = None
    if targets:
        nx, ny, nz = target_counts
        t = block_centroids((nx, ny, nz), (2000.0 / nx, 2000.0 / ny, 500.0 / nz), origin=(0.0, 0.0, -500.0))
    return Workload("C2 ordinary point kriging at scale, 100k samples, 1M targets, <=32 neighbours", x, y, z, value,
                    hole, t, per, shared, search)


def config3(seed=13, radius=150.0, targets=True, target_counts=(200, 200, 100)) -> Workload:
    """C3: 60,000 samples over 1000 x 1000 x 100 m, 90% in 5 Gaussian clusters
    (sigma 60 m), 10% uniform; the C2 model (shared ratios); octant search with
    at most 8 per octant, 4..64; 200 x 200 x 100 blocks of 5 x 5 x 1 m (4M)."""
    rng = np.random.default_rng(seed)
    n = 60_000
    nc = int(0.9 * n)
    centres = rng.uniform([150, 150, -80], [850, 850, -20], (5, 3))
    which = rng.integers(0, 5, nc)
    pc = centres[which] + 60.0 * rng.standard_normal((nc, 3))
    pu = rng.uniform([0, 0, -100], [1000, 1000, 0], (n - nc, 3))
    pts = np.concatenate([pc, pu])
    pts[:, 2] = np.clip(pts[:, 2], -100.0, 0.0)
    pts[:, :2] = np.clip(pts[:, :2], 0.0, 1000.0)
    zf = gaussian_field(pts - pts.mean(0), [("exponential", 0.5, (150.0, 100.0, 40.0)),
                                            ("spherical", 0.4, (400.0, 250.0, 80.0))], 0.1, (30.0, 20.0, 0.0), rng)
    value = lognormal(zf)
    _, shared = _c2_variograms()
    search = _abi.make_search(min_samples=4, max_samples=64, radius=radius, octant=True, max_per_octant=8)
    t = None
    if targets:
        nx, ny, nz = target_counts
        t = block_centroids((nx, ny, nz), (1000.0 / nx, 1000.0 / ny, 100.0 / nz), origin=(0.0, 0.0, -100.0))
    return Workload("C3 uneven batched kriging, 60k clustered samples, 4M targets, octants, 4..64 neighbours",
                    pts[:, 0].copy(), pts[:, 1].copy(), pts[:, 2].copy(), value, None, t, shared, shared, search)


def config4(seed=14) -> Workload:
    """C4: 200,000 samples of the C2 generator (800 holes); LOO with a fixed
    ellipsoid 400/250/80 at (30, 20, 0), at most 32 neighbours."""
    w = config2(seed=seed, holes=(40, 20), targets=False)
    w.name = "C4 differentiable leave-one-out, 200k samples, <=32 neighbours, 12 parameters"
    w.search = _abi.make_search(min_samples=1, max_samples=32, radius=400.0, rotation=(30.0, 20.0, 0.0),
                                ratios=(250.0 / 400.0, 80.0 / 400.0))
    return w


def config5(seed=15, targets=True, target_counts=(200, 200, 50), holes=(40, 40)) -> Workload:
    """C5: block kriging, 10 x 10 x 5 m blocks discretised 4 x 4 x 2, 64
    neighbours; Gaussian covariance (sill 1, range 100 m) with nugget 1e-4, and
    2% of the samples duplicated with a 1 cm jitter (ill-conditioned systems)."""
    rng = np.random.default_rng(seed)
    x, y, z, hole = drillholes(holes[0], holes[1], (2000.0 / holes[0], 2000.0 / holes[1]), 12.5, 250.0, 2.0, rng)
    n = len(x)
    dup = rng.choice(n, int(0.02 * n), replace=False)
    d = rng.standard_normal((len(dup), 3))
    d = 0.01 * d / np.linalg.norm(d, axis=1, keepdims=True)
    x = np.concatenate([x, x[dup] + d[:, 0]])
    y = np.concatenate([y, y[dup] + d[:, 1]])
    z = np.concatenate([z, z[dup] + d[:, 2]])
    hole = np.concatenate([hole, hole[dup]])
    pts = np.stack([x, y, z], 1)
    zf = gaussian_field(pts - pts.mean(0), [("gaussian", 1.0, (100.0, 100.0, 100.0))], 1e-4, (0.0, 0.0, 0.0), rng)
    value = lognormal(zf)
    vg = _abi.make_variogram([("gaussian", 1.0, 100.0)], nugget=1e-4)
    search = _abi.make_search(min_samples=4, max_samples=64, radius=100.0)
    blk = _abi.make_block((10.0, 10.0, 5.0), (4, 4, 2))
    t = None
    if targets:
        nx, ny, nz = target_counts
        t = block_centroids((nx, ny, nz), (2000.0 / nx, 2000.0 / ny, 250.0 / nz), origin=(0.0, 0.0, -250.0))
    return Workload("C5 block kriging, precision-stressed, 2M blocks, 64 neighbours", x, y, z, value, hole, t, vg, vg,
                    search, blk)
