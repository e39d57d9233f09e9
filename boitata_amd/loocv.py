"""Differentiable leave-one-out cross-validation of a nested variogram.

The loss over the held-out samples whose system solves,

    L(theta) = mean_i (z_i - zhat_-i)^2 + mean_i log sigma2_-i,

with fixed neighbour sets (bk_loo_prepare), and its gradient from the
library's adjoint kernel. theta packs, for S structures,

    [nugget, sill_0 .. sill_{S-1},
     major_0, semi_0, minor_0, ..., major_{S-1}, semi_{S-1}, minor_{S-1},
     azimuth, dip, rake]                                     (4 + 4 S values)

ranges in metres and angles in degrees (crates/core/src/rotation.rs:8-16
conventions). ``LooLoss`` is the torch.autograd.Function over it.
"""
from __future__ import annotations

import numpy as np

from . import _abi


def theta_of(variogram: _abi.Variogram) -> np.ndarray:
    S = int(variogram.n_structures)
    th = [variogram.nugget]
    th += [variogram.structures[k].sill for k in range(S)]
    for k in range(S):
        st = variogram.structures[k]
        own = st.ratio_semi > 0 and st.ratio_minor > 0
        rs = st.ratio_semi if own else (variogram.ratio_semi if variogram.has_anisotropy else 1.0)
        rm = st.ratio_minor if own else (variogram.ratio_minor if variogram.has_anisotropy else 1.0)
        th += [st.range, st.range * rs, st.range * rm]
    if variogram.has_anisotropy:
        th += [variogram.azimuth, variogram.dip, variogram.rake]
    else:
        th += [0.0, 0.0, 0.0]
    return np.asarray(th, dtype=np.float64)


def variogram_of(theta, models, params=None) -> _abi.Variogram:
    """The variogram of packed theta; models: one name per structure."""
    th = np.asarray(theta, dtype=np.float64)
    S = len(models)
    if len(th) != 4 + 4 * S:
        raise ValueError(f"theta needs {4 + 4 * S} values for {S} structures")
    structures = []
    for k in range(S):
        major, semi, minor = th[1 + S + 3 * k: 4 + S + 3 * k]
        p = None if params is None else params[k]
        structures.append((models[k], th[1 + k], major, p, (semi / major, minor / major)))
    rot = tuple(th[1 + 4 * S: 4 + 4 * S])
    v = _abi.make_variogram(structures, nugget=th[0], rotation=rot)
    if not v.has_anisotropy:
        # Own ratios still need the (zero) rotation; keep the structures' frames.
        v.has_anisotropy = 1
        v.azimuth = v.dip = v.rake = 0.0
        v.ratio_semi = v.ratio_minor = 1.0
    return v


def loss_and_grad(engine, theta, models, precision=_abi.FP64):
    vg = variogram_of(theta, models)
    loss, grad, used = engine.loo_loss_grad(vg, precision)
    return loss, grad, used


def _function():
    import torch

    class LooLoss(torch.autograd.Function):
        """loss = LooLoss.apply(theta, engine, models); theta a float64 tensor."""

        @staticmethod
        def forward(ctx, theta, engine, models, precision=_abi.FP64):
            th = theta.detach().cpu().numpy().astype(np.float64)
            loss, grad, _ = loss_and_grad(engine, th, models, precision)
            ctx.save_for_backward(torch.from_numpy(grad).to(theta.device, theta.dtype))
            return theta.new_tensor(loss)

        @staticmethod
        def backward(ctx, g):
            (grad,) = ctx.saved_tensors
            return g * grad, None, None, None

    return LooLoss


_LooLoss = None


def loo_loss(theta, engine, models, precision=_abi.FP64):
    """The differentiable loss (a torch scalar) of theta (a torch tensor)."""
    global _LooLoss
    if _LooLoss is None:
        _LooLoss = _function()
    return _LooLoss.apply(theta, engine, models, precision)


def fit_adam(engine, theta0, models, steps=200, lr=0.01, precision=_abi.FP64, scale=None, callback=None):
    """Adam on the leave-one-out loss over theta (in scaled units: each
    parameter divided by `scale`, default |theta0|, so one learning rate fits
    metres, degrees and sills). Returns (theta, history of losses)."""
    import torch

    th0 = np.asarray(theta0, dtype=np.float64)
    sc = np.where(np.abs(th0) > 0, np.abs(th0), 1.0) if scale is None else np.asarray(scale, dtype=np.float64)
    sc_t = torch.tensor(sc, dtype=torch.float64)
    x = torch.tensor(th0 / sc, dtype=torch.float64, requires_grad=True)
    opt = torch.optim.Adam([x], lr=lr)
    hist = []
    for it in range(steps):
        opt.zero_grad()
        loss = loo_loss(x * sc_t, engine, models, precision)
        loss.backward()
        opt.step()
        with torch.no_grad():
            # Keep sills, nugget and ranges positive.
            n_pos = 1 + len(models) * 4
            x[:n_pos].clamp_(min=1e-6)
        hist.append(float(loss.detach()))
        if callback:
            callback(it, hist[-1])
    return (x.detach() * sc_t).numpy(), hist
