"""The leave-one-out gradient at workload scale against FP64 central finite
differences of the oracle's loss (the reference has no gradient; the loss is
mean (z - zhat_-i)^2 + mean log sigma2_-i over the held-out samples, with the
neighbour lists fixed as bk_loo_prepare fixes them):
- all 200,000 held-out systems of the 12-parameter workload (nested
  exponential + spherical with per-structure ranges and three angles), at
  parameters 30% off the generating ones, every component;
- the 204,000 samples of the block workload (Gaussian covariance, 1e-4 nugget,
  1 cm twins: condition numbers ~4e5), the derivatives with respect to the
  nugget and the isotropic range (its three axes moved together).
Both precisions; 1e-5 relative per component with a floor of 1e-3 |grad|."""
import numpy as np
import pytest

import oracle
from boitata_amd import _abi, datasets, loocv

pytestmark = pytest.mark.gpu


def dev(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).to(torch.device("cuda:0"))


class FixedLists:
    def __init__(self, w, search):
        self.d = oracle.Data(w.x, w.y, w.z, w.value, None, w.hole)
        self.P = np.stack([w.x, w.y, w.z], 1)
        self.idx, self.cnt, self.st = oracle.neighbors(self.d, search, None, self.P, left_out=np.arange(w.n))

    def loss(self, vg):
        out = oracle.krige_lists(self.d, vg, self.P, self.idx, self.cnt)
        ok = (out["status"] == 0) & (self.st == 0) & (self.cnt > 0) & (out["variance"] > 0)
        r = self.d.value[ok] - out["value"][ok]
        return float(np.mean(r * r) + np.mean(np.log(out["variance"][ok]))), int(ok.sum())


def check(grad, fd):
    floor = 1e-3 * np.linalg.norm(fd)
    err = np.abs(grad - fd) / np.maximum(np.abs(fd), floor)
    assert err.max() <= 1e-5, (err, grad, fd)
    return float(err.max())


def test_c4_gradient_matches_central_differences(engine):
    w = datasets.config4()
    assert w.n == 200_000
    engine.set_samples(dev(w.x), dev(w.y), dev(w.z), dev(w.value), None, dev(w.hole.astype(np.int32)))
    engine.loo_prepare(w.search, None)
    ref = FixedLists(w, w.search)
    models = ["exponential", "spherical"]
    th_true = loocv.theta_of(w.variogram)
    sign = np.random.default_rng(4).choice([-1.0, 1.0], len(th_true))
    theta = th_true * (1.0 + 0.3 * sign)
    theta[-3:] = th_true[-3:] + 0.3 * sign[-3:] * np.array([30.0, 20.0, 10.0])
    loss0, n0 = ref.loss(loocv.variogram_of(theta, models))
    fd = np.zeros_like(theta)
    for p in range(len(theta)):
        h = 1e-5 * max(abs(theta[p]), 1.0)
        tp, tm = theta.copy(), theta.copy()
        tp[p] += h
        tm[p] -= h
        fd[p] = (ref.loss(loocv.variogram_of(tp, models))[0] - ref.loss(loocv.variogram_of(tm, models))[0]) / (2 * h)
    for prec in (_abi.FP64, _abi.FP32_IR):
        loss, grad, used = engine.loo_loss_grad(loocv.variogram_of(theta, models), prec)
        assert used == n0 and abs(loss - loss0) <= 1e-9 * abs(loss0)
        e = check(grad, fd)
        print(f"C4 gradient vs central differences ({'FP64' if prec == _abi.FP64 else 'FP32+IR'}): max rel {e:.2e}")


def test_c5_range_and_nugget_gradient(engine):
    w = datasets.config5(targets=False)
    engine.set_samples(dev(w.x), dev(w.y), dev(w.z), dev(w.value), None, dev(w.hole.astype(np.int32)))
    engine.loo_prepare(w.search, None)
    ref = FixedLists(w, w.search)
    models = ["gaussian"]
    theta = loocv.theta_of(w.variogram)  # nugget, sill, three ranges, three angles
    assert len(theta) == 8
    loss0, n0 = ref.loss(loocv.variogram_of(theta, models))

    def fd_along(direction, h):
        return (ref.loss(loocv.variogram_of(theta + h * direction, models))[0] -
                ref.loss(loocv.variogram_of(theta - h * direction, models))[0]) / (2 * h)

    e_nug = np.zeros_like(theta)
    e_nug[0] = 1.0
    e_rng = np.zeros_like(theta)
    e_rng[2:5] = 1.0
    fd = np.array([fd_along(e_nug, 1e-4 * theta[0]), fd_along(e_rng, 1e-4 * theta[2])])
    for prec in (_abi.FP64, _abi.FP32_IR):
        loss, grad, used = engine.loo_loss_grad(loocv.variogram_of(theta, models), prec)
        assert used == n0 and abs(loss - loss0) <= 1e-9 * abs(loss0)
        g = np.array([grad[0], grad[2] + grad[3] + grad[4]])
        e = check(g, fd)
        print(f"C5 (nugget, range) gradient vs central differences ({'FP64' if prec == _abi.FP64 else 'FP32+IR'}): "
              f"max rel {e:.2e}; grad {g}, fd {fd}")
