"""Leave-one-out loss and its adjoint gradient on the GPU, against FP64
central finite differences of the oracle's loss and against a dense float64
torch autograd of the same loss (both written here: the reference has no
gradient). Tolerance 1e-5 relative per component."""
import math

import numpy as np
import pytest

import oracle
from boitata_amd import _abi, datasets, loocv

pytestmark = pytest.mark.gpu

MODELS = ["exponential", "spherical"]


def small_set(seed=3):
    rng = np.random.default_rng(seed)
    x, y, z, hole = datasets.drillholes(4, 4, (100.0, 100.0), 30.0, 120.0, 2.0, rng)
    pts = np.stack([x, y, z], 1)
    f = datasets.gaussian_field(pts - pts.mean(0), [("exponential", 0.5, (150.0, 100.0, 40.0)),
                                                    ("spherical", 0.4, (400.0, 250.0, 80.0))], 0.1,
                                (30.0, 20.0, 0.0), rng)
    return x, y, z, datasets.lognormal(f), hole


THETA = np.array([0.13, 0.42, 0.55, 190.0, 120.0, 35.0, 330.0, 260.0, 95.0, 24.0, 14.0, 6.0])
SEARCH = _abi.make_search(min_samples=1, max_samples=16, radius=400.0, rotation=(30.0, 20.0, 0.0),
                          ratios=(250.0 / 400.0, 80.0 / 400.0))


@pytest.fixture(scope="module")
def setup(engine):
    x, y, z, v, hole = small_set()
    engine.set_samples(x, y, z, v, None, hole)
    engine.loo_prepare(SEARCH, None)
    d = oracle.Data(x, y, z, v, None, hole)
    P = np.stack([x, y, z], 1)
    idx, cnt, st = oracle.neighbors(d, SEARCH, None, P, left_out=np.arange(len(x)))
    return d, P, idx, cnt, st


def oracle_loss(setup, theta):
    d, P, idx, cnt, st = setup
    vg = loocv.variogram_of(theta, MODELS)
    out = oracle.krige_lists(d, vg, P, idx, cnt)
    ok = (out["status"] == 0) & (st == 0) & (cnt > 0) & (out["variance"] > 0)
    r = d.value[ok] - out["value"][ok]
    return float(np.mean(r * r) + np.mean(np.log(out["variance"][ok]))), int(ok.sum())


PRECISIONS = [(_abi.FP64, 1e-10), (_abi.FP32_IR, 1e-9)]


@pytest.mark.parametrize("precision,tol", PRECISIONS)
def test_loss_matches_the_oracle(engine, setup, precision, tol):
    loss, grad, used = engine.loo_loss_grad(loocv.variogram_of(THETA, MODELS), precision)
    ref, n = oracle_loss(setup, THETA)
    assert used == n
    assert abs(loss - ref) <= tol * abs(ref)


@pytest.mark.parametrize("precision", [_abi.FP64, _abi.FP32_IR])
def test_gradient_matches_central_differences(engine, setup, precision):
    _, grad, _ = engine.loo_loss_grad(loocv.variogram_of(THETA, MODELS), precision)
    fd = np.zeros_like(THETA)
    for p in range(len(THETA)):
        h = 1e-5 * max(abs(THETA[p]), 1.0)
        tp, tm = THETA.copy(), THETA.copy()
        tp[p] += h
        tm[p] -= h
        fd[p] = (oracle_loss(setup, tp)[0] - oracle_loss(setup, tm)[0]) / (2 * h)
    floor = 1e-3 * np.linalg.norm(fd)
    err = np.abs(grad - fd) / np.maximum(np.abs(fd), floor)
    assert err.max() <= 1e-5, (err, grad, fd)


def dense_loss_torch(setup, theta_t):
    """The same loss built from scratch in float64 torch on the CPU, solved by
    torch.linalg.solve on each bordered system, differentiated by autograd."""
    import torch

    d, P, idx, cnt, st = setup
    keep = np.nonzero((st == 0) & (cnt > 0))[0]
    n = int(cnt[keep].max())
    S = len(MODELS)
    nug, sills = theta_t[0], theta_t[1:1 + S]
    ranges = theta_t[1 + S:1 + 4 * S].reshape(S, 3)
    az, dip, rk = (theta_t[1 + 4 * S + k] * (math.pi / 180.0) for k in range(3))
    sa, ca, sd, cd, sr, cr = torch.sin(az), torch.cos(az), torch.sin(dip), torch.cos(dip), torch.sin(rk), torch.cos(rk)
    z0, o1 = torch.zeros((), dtype=torch.float64), torch.ones((), dtype=torch.float64)
    r1 = torch.stack([torch.stack([sa, ca, z0]), torch.stack([-ca, sa, z0]), torch.stack([z0, z0, o1])])
    r2 = torch.stack([torch.stack([cd, z0, -sd]), torch.stack([z0, o1, z0]), torch.stack([sd, z0, cd])])
    r3 = torch.stack([torch.stack([o1, z0, z0]), torch.stack([z0, cr, sr]), torch.stack([z0, -sr, cr])])
    R = r3 @ r2 @ r1
    Pt = torch.tensor(P, dtype=torch.float64)
    vals = torch.tensor(d.value, dtype=torch.float64)

    def cov(dv):
        rd = dv @ R.T
        zero = (dv == 0).all(-1)
        out = torch.zeros(dv.shape[:-1], dtype=torch.float64)
        for s, model in enumerate(MODELS):
            u = rd / ranges[s]
            t = torch.sqrt((u * u).sum(-1) + zero.double())
            if model == "exponential":
                g = 1 - torch.exp(-3 * t)
            else:
                g = torch.where(t >= 1, torch.ones_like(t), 1.5 * t - 0.5 * t ** 3)
            out = out + sills[s] * (1 - g)
        c0 = nug + sills.sum()
        return torch.where(zero, c0.expand_as(out), out)

    total_sq, total_log, used = 0.0, 0.0, 0
    for b in range(0, len(keep), 512):
        ids = keep[b:b + 512]
        nn = cnt[ids]
        Nb = len(ids)
        K = torch.zeros((Nb, n + 1, n + 1), dtype=torch.float64)
        rhs = torch.zeros((Nb, n + 1), dtype=torch.float64)
        zs = torch.zeros((Nb, n), dtype=torch.float64)
        mask = torch.zeros((Nb, n), dtype=torch.float64)
        nb_idx = torch.tensor(idx[ids, :n].clip(min=0), dtype=torch.long)
        loc = Pt[nb_idx]
        C = cov(loc[:, :, None, :] - loc[:, None, :, :])
        c = cov(Pt[torch.tensor(ids)][:, None, :] - loc)
        for q, k in enumerate(nn):
            k = int(k)
            Kq = torch.zeros((n + 1, n + 1), dtype=torch.float64)
            Kq = Kq + torch.eye(n + 1, dtype=torch.float64)  # padding rows solve to 0
            Kq[:k, :k] = C[q, :k, :k]
            Kq[:k, :k] = Kq[:k, :k] - torch.diag(torch.diag(C[q, :k, :k])) + torch.diag(
                (nug + sills.sum()).expand(k))
            Kq[k:, k:] = torch.eye(n + 1 - k, dtype=torch.float64)
            Kq[:k, n] = 1.0
            Kq[n, :k] = 1.0
            Kq[n, n] = 0.0
            K[q] = Kq
            rhs[q, :k] = c[q, :k]
            rhs[q, n] = 1.0
            zs[q, :k] = vals[nb_idx[q, :k]]
            mask[q, :k] = 1.0
        w = torch.linalg.solve(K, rhs)
        lam = w[:, :n] * mask
        mu = w[:, n]
        zhat = (lam * zs).sum(1)
        s2 = (nug + sills.sum()) - (lam * rhs[:, :n]).sum(1) - mu
        r = vals[torch.tensor(ids)] - zhat
        total_sq = total_sq + (r * r).sum()
        total_log = total_log + torch.log(s2).sum()
        used += Nb
    return total_sq / used + total_log / used


@pytest.mark.parametrize("precision", [_abi.FP64, _abi.FP32_IR])
def test_gradient_matches_dense_autograd(engine, setup, precision):
    import torch

    _, grad, _ = engine.loo_loss_grad(loocv.variogram_of(THETA, MODELS), precision)
    th = torch.tensor(THETA, dtype=torch.float64, requires_grad=True)
    loss = dense_loss_torch(setup, th)
    loss.backward()
    ref = th.grad.numpy()
    floor = 1e-3 * np.linalg.norm(ref)
    err = np.abs(grad - ref) / np.maximum(np.abs(ref), floor)
    assert err.max() <= 1e-5, (err, grad, ref)


def test_autograd_function_and_adam_reduce_the_loss(engine, setup):
    import torch

    th = torch.tensor(THETA, dtype=torch.float64, requires_grad=True)
    loss = loocv.loo_loss(th, engine, MODELS)
    loss.backward()
    _, grad, _ = engine.loo_loss_grad(loocv.variogram_of(THETA, MODELS))
    assert np.allclose(th.grad.numpy(), grad)
    _, hist = loocv.fit_adam(engine, THETA, MODELS, steps=25, lr=0.02)
    assert hist[-1] < hist[0]


@pytest.mark.parametrize("max_samples", [24, 40])
def test_refined_gradient_on_larger_neighbourhoods(engine, setup, max_samples):
    # The 32 and 64 size classes of the refined kernel against the FP64 one
    # (itself pinned above by finite differences and autograd).
    s = _abi.make_search(min_samples=1, max_samples=max_samples, radius=400.0, rotation=(30.0, 20.0, 0.0),
                         ratios=(250.0 / 400.0, 80.0 / 400.0))
    try:
        engine.loo_prepare(s, None)
        vg = loocv.variogram_of(THETA, MODELS)
        l64, g64, u64 = engine.loo_loss_grad(vg, _abi.FP64)
        l32, g32, u32 = engine.loo_loss_grad(vg, _abi.FP32_IR)
        assert u32 == u64
        assert abs(l32 - l64) <= 1e-9 * abs(l64)
        floor = 1e-3 * np.linalg.norm(g64)
        assert (np.abs(g32 - g64) / np.maximum(np.abs(g64), floor)).max() <= 1e-5
    finally:
        engine.loo_prepare(SEARCH, None)


def test_refined_gradient_hands_ill_conditioned_systems_to_fp64(engine, setup):
    # A Gaussian covariance with a 1e-4 nugget and samples duplicated 1 cm
    # away: most FP32 eliminations meet a pivot below their threshold. The
    # refined form must then take the FP64 gradient for those systems, so the
    # loss runs over the same held-out samples as the FP64 form.
    x, y, z, v, hole = small_set()
    rng = np.random.default_rng(7)
    dup = rng.choice(len(x), 12, replace=False)
    jit = rng.normal(size=(12, 3))
    jit *= 0.01 / np.linalg.norm(jit, axis=1, keepdims=True)
    X, Y, Z = np.r_[x, x[dup] + jit[:, 0]], np.r_[y, y[dup] + jit[:, 1]], np.r_[z, z[dup] + jit[:, 2]]
    V, H = np.r_[v, v[dup]], np.r_[hole, hole[dup]]
    s = _abi.make_search(min_samples=1, max_samples=24, radius=400.0)
    theta = np.array([1e-4, 1.0, 100.0, 100.0, 100.0, 0.0, 0.0, 0.0])
    try:
        engine.set_samples(X, Y, Z, V, None, H)
        engine.loo_prepare(s, None)
        vg = loocv.variogram_of(theta, ["gaussian"])
        l64, g64, u64 = engine.loo_loss_grad(vg, _abi.FP64)
        l32, g32, u32 = engine.loo_loss_grad(vg, _abi.FP32_IR)
        assert u32 == u64 and u64 > 0.9 * len(X)
        assert abs(l32 - l64) <= 1e-9 * abs(l64)
        floor = 1e-3 * np.linalg.norm(g64)
        assert (np.abs(g32 - g64) / np.maximum(np.abs(g64), floor)).max() <= 1e-5
    finally:
        x, y, z, v, hole = small_set()
        engine.set_samples(x, y, z, v, None, hole)
        engine.loo_prepare(SEARCH, None)
