"""Engine: one HIP context of the kriging library, driven through the C ABI.

Arrays may be NumPy arrays (host memory, staged by the library) or torch
tensors on the GPU (passed as device pointers, nothing staged).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi
from ._lib import load


class EngineError(RuntimeError):
    """A call of the C ABI returned an error; `code` is the BK_E_* value."""

    def __init__(self, message, code=0):
        super().__init__(message)
        self.code = code


def _ptr(a):
    """Data pointer of a contiguous NumPy array or torch tensor (None -> NULL)."""
    if a is None:
        return None
    if hasattr(a, "data_ptr"):
        if not a.is_contiguous():
            raise ValueError("tensor must be contiguous")
        return C.c_void_p(a.data_ptr())
    if not a.flags.c_contiguous:
        raise ValueError("array must be C-contiguous")
    return C.c_void_p(a.ctypes.data)


def _f64(a):
    if a is None:
        return None
    if hasattr(a, "data_ptr"):
        import torch

        return a.to(torch.float64).contiguous()
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    if a is None:
        return None
    if hasattr(a, "data_ptr"):
        import torch

        return a.to(torch.int32).contiguous()
    return np.ascontiguousarray(a, dtype=np.int32)


def _i64(a):
    if a is None:
        return None
    if hasattr(a, "data_ptr"):
        import torch

        return a.to(torch.int64).contiguous()
    return np.ascontiguousarray(a, dtype=np.int64)


class Engine:
    """A context on HIP device `device` holding the fitted samples."""

    def __init__(self, device: int = 0):
        self._lib = load()
        h = C.c_void_p()
        rc = self._lib.bk_create(C.byref(h), int(device))
        if rc != 0:
            raise EngineError(f"bk_create failed ({rc}): no usable HIP device {device}")
        self._h = h
        self.n = 0

    def close(self):
        if getattr(self, "_h", None):
            self._lib.bk_destroy(self._h)
            self._h = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            msg = self._lib.bk_last_error(self._h).decode()
            raise EngineError(f"{msg} ({rc})", rc)

    def set_samples(self, x, y, z, value, error_variance=None, hole=None):
        x, y, z, value = _f64(x), _f64(y), _f64(z), _f64(value)
        err = _f64(error_variance)
        hole = _i32(hole)
        self._keep = (x, y, z, value, err, hole)
        self.n = int(len(x))
        self._check(self._lib.bk_set_samples(self._h, self.n, _ptr(x), _ptr(y), _ptr(z), _ptr(value), _ptr(err),
                                             _ptr(hole)))

    def profile(self, enable: bool = True):
        """Start (or stop) timing the stages of each call with HIP events."""
        self._check(self._lib.bk_profile(self._h, 1 if enable else 0))

    def profile_read(self):
        """Milliseconds per stage summed since the last read, and the calls timed."""
        ms = np.zeros(6)
        calls = C.c_int64()
        self._check(self._lib.bk_profile_read(self._h, _ptr(ms), C.byref(calls)))
        return dict(zip(_abi.STAGES, ms.tolist())), calls.value

    def tuning(self, edges=None, max_iter=0, tol=0.0, cell_factor=0.0):
        e = None if edges is None else (C.c_int32 * 3)(*edges)
        self._check(self._lib.bk_set_tuning(self._h, e, int(max_iter), float(tol), float(cell_factor)))

    def search_tuning(self, mode: int):
        """Uncapped searches: 0 one wave per target, 1..8 tiles of that reach, -1 one lane per target."""
        self._check(self._lib.bk_set_search_tuning(self._h, int(mode)))

    def solve_tuning(self, waves: int):
        """Wavefronts per workgroup of the solve kernels (1, 2 or 4)."""
        self._check(self._lib.bk_set_solve_tuning(self._h, int(waves)))

    def refine_limit(self, max_n: int):
        """Largest size bucket (16, 32, 64) FP32+IR solves by refinement; larger take FP64."""
        self._check(self._lib.bk_set_refine_limit(self._h, int(max_n)))

    def graphs(self, enable: bool = True):
        """Repeated krige passes over device arrays run as a captured HIP graph (default on)."""
        self._check(self._lib.bk_set_graphs(self._h, 1 if enable else 0))

    def autotune(self, variogram, search, tx, ty, tz, kind=_abi.ORDINARY, param=0.0, precision=_abi.FP32_IR,
                 sample: int = 65536):
        """Times the tunable settings on an evenly strided sample of the targets
        and keeps the fastest in this context; returns the report as a dict."""
        tx, ty, tz = _f64(tx), _f64(ty), _f64(tz)
        m = int(len(tx))
        step = max(m // max(int(sample), 1), 1)
        if step > 1:
            tx, ty, tz = (v[::step].contiguous() if hasattr(v, "data_ptr") else np.ascontiguousarray(v[::step])
                          for v in (tx, ty, tz))
        rep = np.zeros(len(_abi.AUTOTUNE_REPORT))
        self._check(self._lib.bk_autotune(self._h, C.byref(variogram), C.byref(search), int(kind), float(param),
                                          int(precision), int(len(tx)), _ptr(tx), _ptr(ty), _ptr(tz), _ptr(rep)))
        return dict(zip(_abi.AUTOTUNE_REPORT, rep.tolist()))

    def stat(self, which: int = _abi.STAT_SEARCH_SPILL) -> int:
        v = C.c_int64()
        self._check(self._lib.bk_get_stat(self._h, int(which), C.byref(v)))
        return v.value

    def neighbors(self, search: _abi.Search, variogram: _abi.Variogram | None, tx, ty, tz, left_out=None):
        tx, ty, tz = _f64(tx), _f64(ty), _f64(tz)
        left_out = _i64(left_out)
        m = int(len(tx))
        idx = np.full((m, int(search.max_samples)), -1, dtype=np.int32)
        count = np.zeros(m, dtype=np.int32)
        status = np.zeros(m, dtype=np.int32)
        vg = C.byref(variogram) if variogram is not None else None
        self._check(self._lib.bk_neighbors(self._h, C.byref(search), vg, m, _ptr(tx), _ptr(ty), _ptr(tz),
                                           _ptr(left_out), _ptr(idx), _ptr(count), _ptr(status)))
        return idx, count, status

    def _outputs(self, m, like):
        if like is not None and hasattr(like, "data_ptr") and like.is_cuda:
            import torch

            dev = like.device
            f = lambda: torch.empty(m, dtype=torch.float64, device=dev)  # noqa: E731
            i = lambda: torch.empty(m, dtype=torch.int32, device=dev)  # noqa: E731
        else:
            f = lambda: np.empty(m, dtype=np.float64)  # noqa: E731
            i = lambda: np.empty(m, dtype=np.int32)  # noqa: E731
        return f, i

    def krige(self, variogram, search, tx, ty, tz, kind=_abi.ORDINARY, param=0.0, precision=_abi.FP64, out=None):
        """Point kriging of the targets. `out` (a dict as returned, e.g. by an
        earlier call) receives the results in place: with device arrays and
        the same arguments, a repeated call then runs as a captured graph."""
        tx, ty, tz = _f64(tx), _f64(ty), _f64(tz)
        m = int(len(tx))
        if out is not None:
            value, variance, lagrange, n_used, status = (out[k] for k in ("value", "variance", "lagrange", "n_used",
                                                                           "status"))
            for a, dt in ((value, "float64"), (variance, "float64"), (lagrange, "float64"), (n_used, "int32"),
                          (status, "int32")):
                if len(a) != m or str(a.dtype).rsplit(".", 1)[-1] != dt:
                    raise ValueError(f"out arrays must hold {m} values of the returned dtypes")
        else:
            f, i = self._outputs(m, tx)
            value, variance, lagrange, n_used, status = f(), f(), f(), i(), i()
        self._check(self._lib.bk_krige(self._h, C.byref(variogram), C.byref(search), int(kind), float(param),
                                       int(precision), m, _ptr(tx), _ptr(ty), _ptr(tz), _ptr(value), _ptr(variance),
                                       _ptr(lagrange), _ptr(n_used), _ptr(status)))
        return dict(value=value, variance=variance, lagrange=lagrange, n_used=n_used, status=status)

    def krige_noalloc(self, variogram, search, tx, ty, tz, kind=_abi.ORDINARY, param=0.0, precision=_abi.FP64):
        """bk_krige_noalloc: as krige, with the pass's scratch in the library's
        static device array (no allocation by the library). Targets must be
        device tensors; at most 65536 targets, 262144 samples, 64 neighbours."""
        tx, ty, tz = _f64(tx), _f64(ty), _f64(tz)
        m = int(len(tx))
        f, i = self._outputs(m, tx)
        value, variance, lagrange, n_used, status = f(), f(), f(), i(), i()
        self._check(self._lib.bk_krige_noalloc(self._h, C.byref(variogram), C.byref(search), int(kind), float(param),
                                               int(precision), m, _ptr(tx), _ptr(ty), _ptr(tz), _ptr(value),
                                               _ptr(variance), _ptr(lagrange), _ptr(n_used), _ptr(status)))
        return dict(value=value, variance=variance, lagrange=lagrange, n_used=n_used, status=status)

    def block_krige(self, variogram, search, block, tx, ty, tz, kind=_abi.ORDINARY, param=0.0,
                    precision=_abi.FP64):
        tx, ty, tz = _f64(tx), _f64(ty), _f64(tz)
        m = int(len(tx))
        f, i = self._outputs(m, tx)
        value, variance, lagrange, support, n_used, status = f(), f(), f(), f(), i(), i()
        self._check(self._lib.bk_block_krige(self._h, C.byref(variogram), C.byref(search), int(kind), float(param),
                                             C.byref(block), int(precision), m, _ptr(tx), _ptr(ty), _ptr(tz),
                                             _ptr(value), _ptr(variance), _ptr(lagrange), _ptr(support),
                                             _ptr(n_used), _ptr(status)))
        return dict(value=value, variance=variance, lagrange=lagrange, support=support, n_used=n_used,
                    status=status)

    def loo_prepare(self, search, variogram=None):
        """Fixes the leave-one-out neighbour sets of every sample for loo_loss_grad."""
        vg = C.byref(variogram) if variogram is not None else None
        self._check(self._lib.bk_loo_prepare(self._h, C.byref(search), vg))

    def loo_loss_grad(self, variogram, precision=_abi.FP64):
        """Leave-one-out loss and its gradient (4 + 4 S values) on the fixed sets."""
        S = int(variogram.n_structures)
        grad = np.zeros(4 + 4 * S)
        loss = C.c_double()
        used = C.c_int64()
        self._check(self._lib.bk_loo_loss_grad(self._h, C.byref(variogram), int(precision), C.byref(loss),
                                               _ptr(grad), C.byref(used)))
        return loss.value, grad, used.value

    def loo(self, variogram, search, which=None, kind=_abi.ORDINARY, param=0.0, precision=_abi.FP64):
        which = _i64(which)
        m = self.n if which is None else int(len(which))
        value = np.empty(m, dtype=np.float64)
        variance = np.empty(m, dtype=np.float64)
        status = np.empty(m, dtype=np.int32)
        self._check(self._lib.bk_loo(self._h, C.byref(variogram), C.byref(search), int(kind), float(param),
                                     int(precision), m, _ptr(which), _ptr(value), _ptr(variance), _ptr(status)))
        return dict(value=value, variance=variance, status=status)

    def block_loo(self, variogram, search, block, which=None, kind=_abi.ORDINARY, param=0.0, precision=_abi.FP64):
        """bk_block_loo: leave-one-out block kriging at the samples `which` (all by default)."""
        which = _i64(which)
        m = self.n if which is None else int(len(which))
        value = np.empty(m, dtype=np.float64)
        variance = np.empty(m, dtype=np.float64)
        status = np.empty(m, dtype=np.int32)
        self._check(self._lib.bk_block_loo(self._h, C.byref(variogram), C.byref(search), int(kind), float(param),
                                           C.byref(block), int(precision), m, _ptr(which), _ptr(value),
                                           _ptr(variance), _ptr(status)))
        return dict(value=value, variance=variance, status=status)

    def krige_weights(self, variogram, search, tx, ty, tz, kind=_abi.ORDINARY, param=0.0, block=None,
                      precision=_abi.FP64):
        """bk_krige_weights (weights_many): per target the samples used, in the
        estimator's order, and their kriging weights; rows of S = min(max_samples,
        samples) entries, the first count[t] of them valid."""
        tx, ty, tz = _f64(tx), _f64(ty), _f64(tz)
        m = int(len(tx))
        S = int(max(min(search.max_samples, self.n), 1))
        idx = np.zeros((m, S), dtype=np.int32)
        weights = np.zeros((m, S), dtype=np.float64)
        count = np.zeros(m, dtype=np.int32)
        status = np.zeros(m, dtype=np.int32)
        self._check(self._lib.bk_krige_weights(self._h, C.byref(variogram), C.byref(search), int(kind), float(param),
                                               C.byref(block) if block is not None else None, int(precision), m,
                                               _ptr(tx), _ptr(ty), _ptr(tz), _ptr(idx), _ptr(weights), _ptr(count),
                                               _ptr(status)))
        return dict(idx=idx, weights=weights, count=count, status=status)

    def sgs(self, variogram, search, gx, gy, gz, path, z, precision=_abi.FP64):
        """bk_sgs: sequential Gaussian simulation of normal scores at the grid
        nodes (gx, gy, gz), visited in the order `path`, the k-th node of the
        path not on a datum drawing z[k]. The samples set on the engine are the
        conditioning data with their scores as values. Returns per node the
        score and its source (>= 0 the datum it lies on, -1 kriged, -2 drawn
        from the marginal, -3 failed)."""
        gx, gy, gz = _f64(gx), _f64(gy), _f64(gz)
        path, z = _i64(path), _f64(z)
        m = int(len(gx))
        if len(gy) != m or len(gz) != m or len(path) != m or len(z) != m:
            raise ValueError("one coordinate triple, path entry and draw per node")
        if hasattr(gx, "data_ptr") and gx.is_cuda:
            import torch

            score = torch.empty(m, dtype=torch.float64, device=gx.device)
            source = torch.empty(m, dtype=torch.int32, device=gx.device)
        else:
            score = np.empty(m, dtype=np.float64)
            source = np.empty(m, dtype=np.int32)
        self._check(self._lib.bk_sgs(self._h, C.byref(variogram), C.byref(search), int(precision), m, _ptr(gx),
                                     _ptr(gy), _ptr(gz), _ptr(path), _ptr(z), _ptr(score), _ptr(source)))
        return dict(score=score, source=source)

    def _mik_outputs(self, m, k, nc, nq, like):
        """Output arrays of multiple indicator kriging: entry-major (rows, m)."""
        if like is not None and hasattr(like, "data_ptr") and like.is_cuda:
            import torch

            dev = like.device
            f = lambda *s: torch.empty(*s, dtype=torch.float64, device=dev)  # noqa: E731
            i = lambda *s: torch.empty(*s, dtype=torch.int32, device=dev)  # noqa: E731
        else:
            f = lambda *s: np.empty(s, dtype=np.float64)  # noqa: E731
            i = lambda *s: np.empty(s, dtype=np.int32)  # noqa: E731
        out = dict(raw=f(k, m), cdf=f(k, m), correction=f(m), mean=f(m), variance=f(m),
                   probability_above=f(nc, m), mean_above=f(nc, m), quantile_values=f(nq, m),
                   violations=i(m), n_used=i(m), status=i(m))
        ptrs = [_ptr(out[name]) if name in out else None for name, _ in _abi.MikOut._fields_]
        return out, _abi.MikOut(*[p.value if p is not None else None for p in ptrs])

    def mik_predict(self, mik, thresholds, variograms, search, tx, ty, tz, weights=None, block=None, cutoffs=(),
                    quantiles=(), precision=_abi.FP64):
        """Multiple indicator kriging of the targets (one search pass):
        conditional distributions, outputs (rows, m) as documented for bk_mik_out."""
        tx, ty, tz = _f64(tx), _f64(ty), _f64(tz)
        m = int(len(tx))
        thr = np.ascontiguousarray(thresholds, dtype=np.float64)
        vs = list(variograms) if isinstance(variograms, (list, tuple)) else [variograms]
        vgs = (_abi.Variogram * len(vs))(*vs)
        w = _f64(weights)
        cut = np.ascontiguousarray(cutoffs, dtype=np.float64)
        qs = np.ascontiguousarray(quantiles, dtype=np.float64)
        out, o = self._mik_outputs(m, len(thr), len(cut), len(qs), tx)
        self._check(self._lib.bk_mik_predict(self._h, C.byref(mik), _ptr(thr), C.cast(vgs, C.c_void_p), _ptr(w),
                                             C.byref(search), C.byref(block) if block is not None else None,
                                             int(precision), m, _ptr(tx), _ptr(ty), _ptr(tz), len(cut), _ptr(cut),
                                             len(qs), _ptr(qs), C.byref(o)))
        return out

    def mik_cross_validate(self, mik, thresholds, variograms, search, weights=None, which=None,
                           precision=_abi.FP64):
        """Leave-one-out multiple indicator kriging at the samples `which`
        (None: all); adds `pit`, F*(z) at each sample's own value."""
        which = _i64(which)
        m = self.n if which is None else int(len(which))
        thr = np.ascontiguousarray(thresholds, dtype=np.float64)
        vs = list(variograms) if isinstance(variograms, (list, tuple)) else [variograms]
        vgs = (_abi.Variogram * len(vs))(*vs)
        w = _f64(weights)
        out, o = self._mik_outputs(m, len(thr), 0, 0, None)
        pit = np.empty(m, dtype=np.float64)
        self._check(self._lib.bk_mik_cross_validate(self._h, C.byref(mik), _ptr(thr), C.cast(vgs, C.c_void_p),
                                                    _ptr(w), C.byref(search), int(precision), m, _ptr(which),
                                                    _ptr(pit), C.byref(o)))
        out["pit"] = pit
        return out

    def krige_diagnostics(self, variogram, search, tx, ty, tz, kind=_abi.ORDINARY, param=0.0, block=None,
                          precision=_abi.FP64):
        """A kriging pass (point, or block when `block` is given) with the
        per-target diagnostics of bk_diag_out."""
        tx, ty, tz = _f64(tx), _f64(ty), _f64(tz)
        m = int(len(tx))
        f, i = self._outputs(m, tx)
        out = {k: f() for k in ("value", "variance", "lagrange", "support_variance", "negative_weight_sum",
                                "mean_distance")}
        out.update({k: i() for k in ("n_used", "status", "n_holes", "target_met")})
        o = _abi.DiagOut(*[_ptr(out[name]).value for name, _ in _abi.DiagOut._fields_])
        self._check(self._lib.bk_krige_diagnostics(self._h, C.byref(variogram), C.byref(search), int(kind),
                                                   float(param), C.byref(block) if block is not None else None,
                                                   int(precision), m, _ptr(tx), _ptr(ty), _ptr(tz), C.byref(o)))
        return out
