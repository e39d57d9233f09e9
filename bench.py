#!/usr/bin/env python
"""Kriged targets per second of the MI355X engine on the 1M-target workload,
with every other workload of the estimation path reported beside it.

Headline workload (synthetic, seeded): 100,000 drill-hole composites (400
holes, 2 m composites to -500 m), nested variogram nugget 0.1 + exponential
(0.5; 150/100/40 m) + spherical (0.4; 400/250/80 m) rotated (30, 20, 0),
ellipsoid search 400/250/80 m with at most 32 neighbours, 200 x 200 x 25 block
centroids (1,000,000 targets). One step = one pass of the hot path over all
targets with samples and targets resident in HBM: neighbour search,
bucketing, batched solves and the deferred systems. The sample grid is built
once per sample set and kept across passes (the reference builds its search
tree per predict call; the device build takes about 0.1 ms). The headline
`value` is the FP32 + FP64-refinement form; the plain FP64 form runs beside it
for the speed-up and the error of the refined form against it.

The other workloads, each a leg of the report (`configs`):
- C1: 5,000 samples, 50,000 targets, at most 24 neighbours (CPU-runnable case);
- C3: 4,000,000 targets among 60,000 clustered samples, octant search
  (8 per octant), 4-64 neighbours in size buckets;
- C4: leave-one-out cross-validation gradient on 200,000 samples, 12
  variogram parameters, 200 Adam steps;
- C5: 2,000,000 blocks discretised 4 x 4 x 2, 64 neighbours, Gaussian
  covariance with a 1e-4 nugget and 1 cm twins (ill-conditioned systems);
- MIK: multiple indicator kriging of the headline targets, nine thresholds.

Every leg reports the GPU rate of both forms, the stage times from HIP events
on the library's stream, the refined form's sweeps and hand-backs, errors of
the refined form against FP64 (max, p50, p99), and against the oracle (the
C++ restatement of the reference's CPU path; the reference itself cannot be
built here) on a strided sample of the targets, which is also timed as the CPU
baseline on every host thread and on one, with the neighbour-set mismatch
count on that sample.

Prints one JSON line. `roofline` is for the kernel with the most time in the
headline form: its algorithmic flops over its launch time from HIP events,
against the FP32 peak (the refined kernels compute in FP32 on the VALU and the
matrix cores; FP64 only for residuals).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6   # MI355X FP64 (vector and matrix)
PEAK_FP32_TFLOPS = 157.3  # MI355X FP32 (vector and matrix)
PEAK_HBM_GBS = 8000.0

# Per-structure covariance cost in flops (a transcendental counts as one).
SHAPE_COST = {0: 8, 1: 6, 2: 7, 3: 14, 4: 11, 5: 12, 6: 6, 7: 8, 8: 6}

# The solve kernel of a size bucket (upper edge) in each form.
KERNEL_FP32IR = {16: "k_krige32<16>", 32: "k_krige32m", 64: "k_krige64m"}
KERNEL_FP64 = {16: "k_krige64<16>", 32: "k_krige64<32>", 64: "k_krige64<64>"}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def flops_per_system(n, shape_sum, simple=False):
    """Algorithmic flops of one point system of true size n: Cholesky
    n^3/3 + n^2/2 + n/6, two solve pairs 4 n^2, covariances of n(n+1)/2 pairs
    at 11 + sum(c_s) each, 26 n + 5 for projection, Schur scalar, weights,
    value and variance. Simple kriging drops one solve pair."""
    n = np.asarray(n, dtype=np.float64)
    f = n ** 3 / 3 + n ** 2 / 2 + n / 6 + 4 * n ** 2 + n * (n + 1) / 2 * (11 + shape_sum) + 26 * n + 5
    if simple:
        f = f - 2 * n ** 2 - 4 * n
    return np.where(n > 0, f, 0.0)


def block_flops_per_system(n, points, shape_sum):
    """A block system adds the n x P point-to-block covariances. The P x P
    block-to-block average is one number per call for a stationary model and
    is not credited per block."""
    n = np.asarray(n, dtype=np.float64)
    return flops_per_system(n, shape_sum) + np.where(n > 0, n * points * (11 + shape_sum), 0.0)


def search_bytes_per_target(n):
    """Least HBM traffic of one search: the target (24 B), the selected
    samples' projected coordinates and index (32 B each), the list written (4 B each)."""
    return 24.0 + 36.0 * np.asarray(n, dtype=np.float64)


def measured_traffic(key):
    """HBM bytes per launch recorded by the PMC summary in profiles/ (the
    FETCH_SIZE and WRITE_SIZE passes scaled by the copy calibration), or None."""
    path = os.path.join(ROOT, "profiles", "k_krige32m_c2_pmc.md")
    try:
        with open(path) as f:
            for ln in f:
                if ln.startswith(key + ":"):
                    return float(ln.split(":", 1)[1])
    except OSError:
        pass
    return None


def total_sill(vg):
    return vg.nugget + sum(vg.structures[k].sill for k in range(vg.n_structures))


def warm_up(run, seconds=0.3):
    """Call `run` until about `seconds` have passed (once at least), so that
    a timed loop starts on a clocked-up GPU after a stretch of CPU work."""
    import torch

    t0 = time.perf_counter()
    while True:
        run()
        torch.cuda.synchronize()
        if time.perf_counter() - t0 >= seconds:
            return


def rel_stats(x, y, floor):
    """Max, median and 99th percentile of |x - y| / max(|y|, floor)."""
    x, y = np.asarray(x), np.asarray(y)
    if len(x) == 0:
        return None
    r = np.abs(x - y) / np.maximum(np.abs(y), floor)
    return {"max": float(r.max()), "p50": float(np.percentile(r, 50)), "p99": float(np.percentile(r, 99))}


def strided(m, k):
    """k evenly strided indices of range(m) (all of them when k >= m)."""
    k = int(max(1, min(m, k)))
    return np.arange(0, m, max(m // k, 1), dtype=np.int64)[:k]


def host(a):
    return a.cpu().numpy() if hasattr(a, "cpu") else np.asarray(a)


class Leg:
    """One workload on a context of its own: the timed passes of both forms,
    the stage times, the error of the refined form against FP64, and the
    oracle's CPU baseline and error on a strided sample of the targets."""

    def __init__(self, name, w, args, threads, dev, tuned_sample=65536):
        import torch

        from boitata_amd.engine import Engine

        self.name, self.w, self.args, self.threads = name, w, args, threads
        self.tt = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
        self.e = Engine(0)
        holes = None if w.hole is None else self.tt(w.hole.astype(np.int32))
        self.e.set_samples(self.tt(w.x), self.tt(w.y), self.tt(w.z), self.tt(w.value), None, holes)
        self.T = w.targets
        self.m = len(self.T)
        self.tdev = tuple(self.tt(self.T[:, k]) for k in range(3))
        self.block = w.block
        self.outs = {}
        self.tuned = None
        if not args.no_autotune and self.block is None:
            from boitata_amd import _abi

            t0 = time.perf_counter()
            self.tuned = self.e.autotune(w.variogram, w.search, *self.tdev, precision=_abi.FP32_IR,
                                         sample=min(tuned_sample, self.m))
            self.tuned["seconds"] = time.perf_counter() - t0

    def run(self, prec):
        vg, s = self.w.variogram, self.w.search
        if self.block is None:
            # Each form writes into its own output arrays, so a repeated pass
            # has the same arguments and the library runs it as a captured graph.
            self.outs[prec] = self.e.krige(vg, s, *self.tdev, precision=prec, out=self.outs.get(prec))
        else:
            self.outs[prec] = self.e.block_krige(vg, s, self.block, *self.tdev, precision=prec)
        return self.outs[prec]

    def gpu(self, steps, warmup):
        import torch

        from boitata_amd import _abi

        for _ in range(max(warmup, 1)):
            self.run(_abi.FP32_IR)
            self.run(_abi.FP64)
        dt = {}
        for prec in (_abi.FP32_IR, _abi.FP64):
            warm_up(lambda: self.run(prec))
            t0 = time.perf_counter()
            for _ in range(steps):
                self.run(prec)
            torch.cuda.synchronize()
            dt[prec] = (time.perf_counter() - t0) / steps
        self.dt32, self.dt64 = dt[_abi.FP32_IR], dt[_abi.FP64]
        # Refinement counters of the last refined pass.
        self.run(_abi.FP32_IR)
        refined = self.e.stat(_abi.STAT_REFINED)
        sweeps = self.e.stat(_abi.STAT_REFINE_SWEEPS)
        redo = self.e.stat(_abi.STAT_REFINE_REDO)
        self.refine = {"refined": refined, "sweeps_per_system": sweeps / refined if refined else None,
                       "handed_to_fp64": redo}
        # Stage times: the same passes again with HIP events on the library's
        # stream (a profiled pass runs its launches one by one, not as a graph).
        self.stages = {}
        for prec in (_abi.FP32_IR, _abi.FP64):
            self.e.profile(True)
            for _ in range(steps):
                self.run(prec)
            st, calls = self.e.profile_read()
            self.e.profile(False)
            self.stages[prec] = {k: v / max(calls, 1) for k, v in st.items()}
        # The cell grid is built once per set of samples (and per search frame
        # and cell factor) and reused by every pass; the timed passes above
        # only check it. Its cost: the samples set again, then one pass.
        w = self.w
        holes = None if w.hole is None else self.tt(w.hole.astype(np.int32))
        xs = (self.tt(w.x), self.tt(w.y), self.tt(w.z), self.tt(w.value))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        self.e.set_samples(*xs, None, holes)
        torch.cuda.synchronize()
        t_set = time.perf_counter() - t0
        self.e.profile(True)
        self.run(_abi.FP32_IR)
        st, _ = self.e.profile_read()
        self.e.profile(False)
        self.grid_build = {"set_samples_ms": t_set * 1e3, "grid_build_ms": st["grid"],
                           "note": "once per set of samples, search frame and cell factor; not in the timed passes"}
        a32, a64 = self.outs[_abi.FP32_IR], self.outs[_abi.FP64]
        self.s32, self.s64 = host(a32["status"]), host(a64["status"])
        self.h32 = {k: host(a32[k]) for k in ("value", "variance")}
        self.h64 = {k: host(a64[k]) for k in ("value", "variance")}
        self.n_used = host(a64["n_used"])

    def report(self):
        from boitata_amd import _abi

        vg = self.w.variogram
        c0 = total_sill(vg)
        floor = 1e-3 * c0
        ok = (self.s32 == 0) & (self.s64 == 0)
        est = self.s64 == 0
        nu = self.n_used[est]
        shape_sum = sum(SHAPE_COST[vg.structures[k].model] for k in range(vg.n_structures))
        if self.block is None:
            f = flops_per_system(nu, shape_sum)
        else:
            pts = int(self.block.nx) * int(self.block.ny) * int(self.block.nz)
            f = block_flops_per_system(nu, pts, shape_sum)
        flops = float(f.sum())
        sbytes = float(search_bytes_per_target(self.n_used).sum())
        st32, st64 = self.stages[_abi.FP32_IR], self.stages[_abi.FP64]
        # Systems handed back to FP64 are not credited to the refined kernels.
        redo = self.refine["handed_to_fp64"]
        f32 = flops * (1.0 - redo / max(int(est.sum()), 1))
        buckets = sorted({16 if n <= 16 else (32 if n <= 32 else 64) for n in np.unique(nu)})
        solve32 = f32 / (st32["solve"] * 1e-3) / 1e12 if st32["solve"] > 0 else None
        solve64 = flops / (st64["solve"] * 1e-3) / 1e12 if st64["solve"] > 0 else None
        search = sbytes / (st32["search"] * 1e-3) / 1e9 if st32["search"] > 0 else None
        return {
            "workload": self.w.name, "targets": self.m, "samples": self.w.n,
            "support": "point" if self.block is None else
                       f"block {int(self.block.nx)}x{int(self.block.ny)}x{int(self.block.nz)}",
            "value": self.m / self.dt32, "unit": "targets/s", "ms_per_step": self.dt32 * 1e3,
            "fp64": {"value": self.m / self.dt64, "ms_per_step": self.dt64 * 1e3},
            "speedup_fp32ir_over_fp64": self.dt64 / self.dt32,
            "stages_ms_fp32ir": st32, "stages_ms_fp64": st64,
            "refinement": self.refine, "grid": self.grid_build,
            "rel_err_fp32ir_vs_fp64": {k: rel_stats(self.h32[k][ok], self.h64[k][ok], floor)
                                       for k in ("value", "variance")},
            "rel_err_floor": floor,
            "status_equal_fp32ir_fp64": bool((self.s32 == self.s64).all()),
            "estimated": int(est.sum()),
            "n_used": {"min": int(nu.min()) if len(nu) else 0, "max": int(nu.max()) if len(nu) else 0,
                       "le16": int((nu <= 16).sum()), "le32": int(((nu > 16) & (nu <= 32)).sum()),
                       "le64": int(((nu > 32) & (nu <= 64)).sum()), "gt64": int((nu > 64).sum())},
            "kernels": {"fp32ir": [KERNEL_FP32IR[b] for b in buckets], "fp64": [KERNEL_FP64[b] for b in buckets]},
            "solve_fp32ir": {"tflops": solve32, "frac_fp32_peak": solve32 / PEAK_FP32_TFLOPS if solve32 else None},
            "solve_fp64": {"tflops": solve64, "frac_fp64_peak": solve64 / PEAK_FP64_TFLOPS if solve64 else None},
            "search": {"gbs": search, "frac_hbm_peak": search / PEAK_HBM_GBS if search else None},
            "autotune": self.tuned,
        }

    def oracle_leg(self, seconds, neighbour_sample=20000, kappa_sample=0):
        """Times the oracle on a strided sample (all host threads, then one
        thread on a smaller sample) and compares both GPU forms and the GPU
        neighbour lists with it there."""
        import oracle
        from boitata_amd import _abi

        w = self.w
        d = oracle.Data(w.x, w.y, w.z, w.value, None, w.hole)

        def solve(idx, threads):
            P = self.T[idx]
            if self.block is None:
                return oracle.krige(d, w.variogram, w.search, P, threads=threads)
            return oracle.block_krige(d, w.variogram, w.search, self.block, P, threads=threads)

        probe = strided(self.m, 1000)
        t0 = time.perf_counter()
        solve(probe, self.threads)
        rate = len(probe) / (time.perf_counter() - t0)
        idx = strided(self.m, rate * seconds)
        # Whole passes over the sample until about `seconds` of work.
        passes, t0 = 0, time.perf_counter()
        while True:
            o = solve(idx, self.threads)
            passes += 1
            dt = time.perf_counter() - t0
            if dt >= seconds / 2 or passes >= 50:
                break
        idx1 = strided(self.m, rate / self.threads * seconds / 3)
        t0 = time.perf_counter()
        solve(idx1, 1)
        dt1 = time.perf_counter() - t0
        stride = max(self.m // len(idx), 1)
        what = "neighbour search + FP64 " + ("ordinary kriging" if self.block is None else "block kriging (LU)")
        cpu = {"value": passes * len(idx) / dt, "unit": "targets/s", "cores": self.threads, "kind": "port",
               "sample": f"{passes} pass(es) over {len(idx)} of the {self.m} targets (every {stride}th), {what}, "
                         f"{dt:.1f} s",
               "single_thread": {"value": len(idx1) / dt1, "unit": "targets/s", "cores": 1,
                                 "sample": f"{len(idx1)} of the {self.m} targets, {dt1:.1f} s"}}
        floor = 1e-3 * total_sill(w.variogram)
        ok = (o["status"] == 0) & (self.s64[idx] == 0) & (self.s32[idx] == 0)
        err = {}
        for form, h in (("fp64", self.h64), ("fp32ir", self.h32)):
            err[form] = {k: rel_stats(h[k][idx][ok], o[k][ok], floor) for k in ("value", "variance")}
        # Neighbour sets on (a part of) the same sample.
        nidx = idx[strided(len(idx), neighbour_sample)]
        P = self.T[nidx]
        oi, oc, os_ = oracle.neighbors(d, w.search, w.variogram, P, threads=self.threads)
        gi, gc, gs = self.e.neighbors(w.search, w.variogram, P[:, 0], P[:, 1], P[:, 2])
        bad = (oc != gc) | (os_ != gs)
        cols = np.arange(oi.shape[1])[None, :] < oc[:, None]
        bad |= ((oi != gi) & cols).any(1)
        out = {"cpu_baseline": cpu, "oracle_sample": len(idx), "rel_err_vs_oracle": err,
               "status_mismatch_vs_oracle": {"fp64": int((self.s64[idx] != o["status"]).sum()),
                                             "fp32ir": int((self.s32[idx] != o["status"]).sum())},
               "neighbour_sets_compared": len(nidx), "neighbour_set_mismatches": int(bad.sum())}
        if kappa_sample > 0:
            # 2-norm condition numbers of the covariance matrices of a strided
            # subset of the sample's systems.
            kk = nidx[strided(len(nidx), kappa_sample)]
            sel = np.searchsorted(nidx, kk)
            kap = np.array([np.linalg.cond(oracle.system_matrix(d, w.variogram, oi[j, :oc[j]]))
                            for j in sel if oc[j] > 0])
            out["kappa"] = {"systems": len(kap), "p50": float(np.median(kap)),
                            "p99": float(np.percentile(kap, 99)), "max": float(kap.max())}
        return out

    def close(self):
        self.e.close()


def roofline_of(rep, traffic):
    """The roofline object of the headline: the kernel with the most time."""
    st = rep["stages_ms_fp32ir"]
    if st["solve"] >= st["search"]:
        ach = rep["solve_fp32ir"]["tflops"]
        return dict(bound="valu", achieved=ach, peak=PEAK_FP32_TFLOPS, unit="TFLOP/s",
                    frac=ach / PEAK_FP32_TFLOPS, traffic=traffic, kernel=" + ".join(rep["kernels"]["fp32ir"]),
                    note="compute-bound batched solves: F(n) algorithmic flops of every system the refined "
                         "kernels keep, over the solve stage's HIP-event time, against the FP32 peak (the "
                         "kernels compute in FP32 on the VALU, elimination on FP32 MFMA, residuals in FP64)")
    ach = rep["search"]["gbs"]
    return dict(bound="hbm", achieved=ach, peak=PEAK_HBM_GBS, unit="GB/s", frac=ach / PEAK_HBM_GBS,
                traffic=None, kernel="k_knn_wave")


def loo_leg(args, dev, threads):
    """C4: Adam on the leave-one-out loss through the adjoint kernels."""
    import torch

    from boitata_amd import _abi, datasets, loocv
    from boitata_amd.engine import Engine

    tt = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    w4 = datasets.config4()
    models = ["exponential", "spherical"]
    e = Engine(0)
    e.set_samples(tt(w4.x), tt(w4.y), tt(w4.z), tt(w4.value), None, tt(w4.hole.astype(np.int32)))
    t0 = time.perf_counter()
    e.loo_prepare(w4.search, None)
    torch.cuda.synchronize()
    prep = time.perf_counter() - t0
    th_true = loocv.theta_of(w4.variogram)
    sign = np.random.default_rng(4).choice([-1.0, 1.0], len(th_true))
    th0 = th_true * (1.0 + 0.3 * sign)
    th0[-3:] = th_true[-3:] + 0.3 * sign[-3:] * np.array([30.0, 20.0, 10.0])  # angles: degrees off

    def adam_rate(prec, steps):
        loocv.fit_adam(e, th0, models, steps=2, lr=0.02, precision=prec)  # warm-up (first optimizer step)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        th, hist = loocv.fit_adam(e, th0, models, steps=steps, lr=0.02, precision=prec)
        torch.cuda.synchronize()
        return time.perf_counter() - t0, th, hist

    dtl, th, hist = adam_rate(_abi.FP32_IR, args.loo_steps)
    log(f"LOO-CV: {args.loo_steps / dtl:.1f} Adam steps/s on {w4.n} samples, loss {hist[0]:.4f} -> {hist[-1]:.4f}")
    s64 = max(args.loo_steps // 4, 1)
    dtl64, _, _ = adam_rate(_abi.FP64, s64)
    loss, grad, used = loocv.loss_and_grad(e, th, models, _abi.FP32_IR)
    loss64, grad64, _ = loocv.loss_and_grad(e, th, models, _abi.FP64)
    gfloor = 1e-3 * float(np.linalg.norm(grad64))
    # Central differences of the FP64 device loss at the starting parameters
    # (Adam drives the nugget towards 0, where a central step leaves the
    # domain). The device FP64 loss equals the oracle's to 1e-9 relative on
    # these fixed sets; the gradient tests difference the oracle's loss itself.
    fd = np.zeros_like(th0)
    for p in range(len(th0)):
        h = 1e-5 * max(abs(th0[p]), 1.0)
        tp, tm = th0.copy(), th0.copy()
        tp[p] += h
        tm[p] -= h
        fd[p] = (loocv.loss_and_grad(e, tp, models, _abi.FP64)[0] -
                 loocv.loss_and_grad(e, tm, models, _abi.FP64)[0]) / (2 * h)
    ffloor = 1e-3 * float(np.linalg.norm(fd))
    g0_32 = loocv.loss_and_grad(e, th0, models, _abi.FP32_IR)[1]
    g0_64 = loocv.loss_and_grad(e, th0, models, _abi.FP64)[1]

    def rel(g):
        return float(np.max(np.abs(g - fd) / np.maximum(np.abs(fd), ffloor)))

    line = {"workload": w4.name, "value": args.loo_steps / dtl, "unit": "Adam steps/s", "steps": args.loo_steps,
            "ms_per_step": dtl / args.loo_steps * 1e3, "samples": w4.n, "systems_per_step": used,
            "parameters": len(th0), "lr": 0.02, "loss_first": hist[0], "loss_last": hist[-1],
            "neighbour_sets_fixed_s": prep,
            "precision": "FP32 elimination forward with FP64 refinement, FP32 adjoint, FP64 reduction",
            "fp64": {"value": s64 / dtl64, "unit": "Adam steps/s"},
            "speedup_fp32ir_over_fp64": (args.loo_steps / dtl) / (s64 / dtl64),
            "rel_err_vs_fp64": {"loss": abs(loss - loss64) / abs(loss64),
                                "grad": float(np.max(np.abs(grad - grad64) / np.maximum(np.abs(grad64), gfloor)))},
            "grad_rel_err_vs_central_differences": {"fp32ir": rel(g0_32), "fp64": rel(g0_64),
                                                    "at": "starting parameters", "floor": "1e-3 |fd|",
                                                    "step": "1e-5 max(|theta|, 1)"}}
    e.close()
    if threads:
        import oracle

        d = oracle.Data(w4.x, w4.y, w4.z, w4.value, None, w4.hole)
        vg = loocv.variogram_of(th, models)
        which = strided(w4.n, 2000)
        t0 = time.perf_counter()
        oracle.loo(d, vg, w4.search, which=which, threads=threads)
        rate = len(which) / (time.perf_counter() - t0)
        which = strided(w4.n, rate * args.cpu_seconds)
        t0 = time.perf_counter()
        oracle.loo(d, vg, w4.search, which=which, threads=threads)
        dt = time.perf_counter() - t0
        rate = len(which) / dt
        # A CPU optimiser without an adjoint needs 2 P + 1 loss passes per step.
        line["cpu_baseline"] = {
            "value": rate / w4.n / (2 * len(th) + 1), "unit": "Adam steps/s (central differences)",
            "cores": threads, "kind": "port",
            "sample": f"oracle leave-one-out (search + FP64 solve) of {len(which)} of the {w4.n} samples in "
                      f"{dt:.1f} s: {rate:.0f} systems/s; a step of central differences is {2 * len(th) + 1} "
                      "passes over all samples"}
    return line


def mik_leg(args, w, dev, threads):
    """Multiple indicator kriging of the headline targets: the nine deciles of
    the data as thresholds with the one variogram (median indicator kriging),
    three cutoffs and three quantiles; one pass = search, solves writing
    weights, indicator estimates, conditional distributions."""
    import torch

    from boitata_amd import _abi
    from boitata_amd.engine import Engine

    tt = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    e = Engine(0)
    e.set_samples(tt(w.x), tt(w.y), tt(w.z), tt(w.value), None, tt(w.hole.astype(np.int32)))
    T = w.targets
    m = len(T)
    tx, ty, tz = tt(T[:, 0]), tt(T[:, 1]), tt(T[:, 2])
    vg, search = w.variogram, w.search
    thr = np.quantile(w.value, np.linspace(0.1, 0.9, 9))
    mk = _abi.make_mik(thr, 1)
    kw = dict(cutoffs=[0.5, 1.0, 2.0], quantiles=[0.1, 0.5, 0.9])
    mres = {}
    for prec in (_abi.FP32_IR, _abi.FP64):
        warm_up(lambda: e.mik_predict(mk, thr, [vg], search, tx, ty, tz, precision=prec, **kw))
        t0 = time.perf_counter()
        for _ in range(args.mik_steps):
            mres[prec] = e.mik_predict(mk, thr, [vg], search, tx, ty, tz, precision=prec, **kw)
        torch.cuda.synchronize()
        mres[prec]["dt"] = (time.perf_counter() - t0) / args.mik_steps
    a32, a64 = mres[_abi.FP32_IR], mres[_abi.FP64]
    okm = ((a32["status"] == 0) & (a64["status"] == 0)).cpu().numpy()

    def rel(k):
        x, y = a32[k].cpu().numpy()[..., okm], a64[k].cpu().numpy()[..., okm]
        return rel_stats(x.ravel(), y.ravel(), 1e-3) if okm.any() else None

    line = {"value": m / a32["dt"], "unit": "targets/s", "ms_per_step": a32["dt"] * 1e3,
            "steps": args.mik_steps, "thresholds": len(thr), "variograms": 1, "cutoffs": 3, "quantiles": 3,
            "precision": "FP32 elimination weights with FP64 refinement",
            "fp64": {"value": m / a64["dt"], "ms_per_step": a64["dt"] * 1e3},
            "rel_err_fp32ir_vs_fp64": {"cdf": rel("cdf"), "mean": rel("mean"), "variance": rel("variance")},
            "estimated": int(okm.sum())}
    log(f"MIK: {line['value']:.0f} targets/s FP32+IR, {line['fp64']['value']:.0f} FP64")
    e.close()
    if threads:
        import oracle

        d = oracle.Data(w.x, w.y, w.z, w.value, None, w.hole)
        few = T[strided(m, 20000)]
        t0 = time.perf_counter()
        oracle.mik_predict(d, mk, thr, [vg], search, few, threads=threads, **kw)
        rate = len(few) / (time.perf_counter() - t0)
        few = T[strided(m, rate * args.cpu_seconds)]
        t0 = time.perf_counter()
        oracle.mik_predict(d, mk, thr, [vg], search, few, threads=threads, **kw)
        dtm = time.perf_counter() - t0
        line["cpu_baseline"] = dict(value=len(few) / dtm, unit="targets/s", cores=threads, kind="port",
                                    sample=f"{len(few)} of the {m} targets, {dtm:.1f} s")
    return line


def sgs_leg(args, dev, threads):
    """Sequential Gaussian simulation of normal scores: the 100,000 samples of
    the headline workload (rank normal scores of their values), 1,000,000
    nodes of a 100 x 100 x 100 grid over the same volume, the headline's
    ellipsoid search (at most 32 neighbours) and nested model scaled to a unit
    sill, a random path and standard normal draws (numpy, seed 5)."""
    import torch
    from scipy.stats import norm

    from boitata_amd import _abi, datasets
    from boitata_amd.engine import Engine

    w = datasets.config2(targets=False)
    r = np.argsort(np.argsort(w.value, kind="stable"), kind="stable")
    scores = norm.ppf((r + 0.5) / w.n)
    n_side = args.sgs_side
    axes = [np.linspace(lo + (hi - lo) / (2 * n_side), hi - (hi - lo) / (2 * n_side), n_side)
            for lo, hi in ((0.0, 1000.0), (0.0, 1000.0), (-500.0, 0.0))]
    G = np.stack(np.meshgrid(*axes, indexing="ij"), -1).reshape(-1, 3)
    m = len(G)
    rng = np.random.default_rng(5)
    path = rng.permutation(m)
    z = rng.normal(size=m)
    tt = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    e = Engine(0)
    e.set_samples(tt(w.x), tt(w.y), tt(w.z), tt(scores), None, tt(w.hole.astype(np.int32)))
    gx, gy, gz, dpath, dz = tt(G[:, 0]), tt(G[:, 1]), tt(G[:, 2]), tt(path), tt(z)
    res = {}
    for prec in (_abi.FP32_IR, _abi.FP64):
        warm_up(lambda: e.sgs(w.variogram, w.search, gx, gy, gz, dpath, dz, precision=prec))
        t0 = time.perf_counter()
        for _ in range(args.sgs_steps):
            out = e.sgs(w.variogram, w.search, gx, gy, gz, dpath, dz, precision=prec)
        torch.cuda.synchronize()
        res[prec] = ((time.perf_counter() - t0) / args.sgs_steps, out, e.stat(_abi.STAT_SGS_PASSES))
    (dt32, o32, passes), (dt64, o64, _) = res[_abi.FP32_IR], res[_abi.FP64]
    s32, s64 = host(o32["score"]), host(o64["score"])
    src = host(o64["source"])
    line = {"workload": f"SGS in normal scores: {w.n} samples, {m} nodes, <=32 neighbours",
            "value": m / dt32, "unit": "nodes/s", "ms_per_realization": dt32 * 1e3,
            "fp64": {"value": m / dt64, "ms_per_realization": dt64 * 1e3},
            "speedup_fp32ir_over_fp64": dt64 / dt32, "draw_passes": passes,
            "rel_err_fp32ir_vs_fp64": rel_stats(s32, s64, 1e-3),
            "nodes": {"kriged": int((src == -1).sum()), "marginal": int((src == -2).sum()),
                      "on_data": int((src >= 0).sum())},
            "steps": args.sgs_steps}
    log(f"SGS: {line['value']:.0f} nodes/s FP32+IR, {line['fp64']['value']:.0f} FP64, {passes} draw passes")
    e.close()
    if threads:
        import oracle

        # The oracle's plan form (parallel searches and solves, draws in path
        # order) on the same data, nodes, path and draws; its scores are the
        # device's FP64 ones within 1e-9.
        d = oracle.Data(w.x, w.y, w.z, scores, None, w.hole)
        t0 = time.perf_counter()
        cs, _ = oracle.sgs(d, w.variogram, w.search, G, path, z, plan=True, threads=threads)
        dtc = time.perf_counter() - t0
        line["cpu_baseline"] = {"value": m / dtc, "unit": "nodes/s", "cores": threads, "kind": "port",
                                "sample": f"one realization of the same {m} nodes (per-node plans), {dtc:.1f} s"}
        line["rel_err_fp64_vs_oracle"] = rel_stats(s64, cs, 1e-3)
    return line


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines and oracle comparisons")
    ap.add_argument("--cpu-seconds", type=float, default=8.0, help="all-thread oracle time per leg")
    ap.add_argument("--targets", type=int, default=0, help="headline: use the first N targets only (0: all)")
    ap.add_argument("--loo-steps", type=int, default=200, help="Adam steps of the C4 leg (0: skip)")
    ap.add_argument("--no-autotune", action="store_true", help="keep the library's default tuning")
    ap.add_argument("--mik-steps", type=int, default=5, help="timed passes of the MIK leg (0: skip)")
    ap.add_argument("--config-steps", type=int, default=2,
                    help="timed passes of the C1, C3 and C5 legs (0: skip them)")
    ap.add_argument("--legs", default="c1,c3,c4,c5,mik,sgs", help="comma-separated legs beside the headline")
    ap.add_argument("--sgs-side", type=int, default=100, help="SGS leg: nodes per grid side")
    ap.add_argument("--sgs-steps", type=int, default=2, help="SGS leg: timed realizations per form")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="finish with a device-to-device copy of this many MiB (byte calibration of PMC runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the engine shards targets across GPUs "
              "by design but no multi-GPU figure is produced", file=sys.stderr)
        sys.exit(2)

    import torch

    from boitata_amd import _abi, datasets

    if not torch.cuda.is_available():
        print("no GPU", file=sys.stderr)
        sys.exit(1)
    dev = torch.device("cuda:0")
    legs = {s.strip() for s in args.legs.split(",") if s.strip()}
    threads = 0 if args.no_cpu else int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))

    w = datasets.config2()
    if args.targets > 0:
        w.targets = w.targets[: args.targets]
    m = len(w.targets)
    log(f"headline: {w.n} samples, {m} targets")
    head = Leg("C2", w, args, threads, dev)
    if head.tuned:
        log(f"autotune: {head.tuned}")
    head.gpu(args.steps, args.warmup)
    log(f"FP32+IR {head.dt32 * 1e3:.2f} ms/step, FP64 {head.dt64 * 1e3:.2f} ms/step")
    rep = head.report()
    graph_launches = head.e.stat(_abi.STAT_GRAPH_LAUNCHES)
    if threads:
        rep.update(head.oracle_leg(args.cpu_seconds))
        log(f"C2 CPU oracle {rep['cpu_baseline']['value']:.0f} targets/s on {threads} threads; "
            f"neighbour mismatches {rep['neighbour_set_mismatches']}")
    head.close()
    traffic = measured_traffic("traffic_bytes_k_krige32m_c2") if m == 1_000_000 else None
    roof = roofline_of(rep, traffic)

    configs = {"C2": rep}
    if args.config_steps > 0:
        for key, gen, kap in (("C1", datasets.config1, 0), ("C3", datasets.config3, 0), ("C5", datasets.config5, 400)):
            if key.lower() not in legs:
                continue
            leg = Leg(key, gen(), args, threads, dev)
            leg.gpu(args.config_steps, 1)
            r = leg.report()
            log(f"{key}: {r['value']:.0f} targets/s FP32+IR, {r['fp64']['value']:.0f} FP64, "
                f"x{r['speedup_fp32ir_over_fp64']:.2f}")
            if threads:
                r.update(leg.oracle_leg(args.cpu_seconds, kappa_sample=kap))
                log(f"{key} CPU oracle {r['cpu_baseline']['value']:.0f} targets/s; neighbour mismatches "
                    f"{r['neighbour_set_mismatches']}")
            configs[key] = r
            leg.close()
            del leg
    if args.loo_steps > 0 and "c4" in legs:
        configs["C4"] = loo_leg(args, dev, threads)
    mik = mik_leg(args, w, dev, threads) if args.mik_steps > 0 and "mik" in legs else None
    sgs = sgs_leg(args, dev, threads) if args.sgs_steps > 0 and "sgs" in legs else None

    other = {"search_c2": rep["search"], "solve_fp64_c2": rep["solve_fp64"]}
    for key in ("C3", "C5"):
        if key in configs:
            other[f"solve_fp32ir_{key.lower()}"] = dict(configs[key]["solve_fp32ir"],
                                                     kernels=configs[key]["kernels"]["fp32ir"])
    line = {
        "metric": "kriged targets/s (FP32+IR vs FP64) at <=1e-6 rel err; LOO-CV grad steps/s",
        "value": m / head.dt32,
        "unit": "targets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head.dt32 * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp64 refinement",
        "data": "synthetic (seeded drill-hole composites, lognormal Gaussian-field values)",
        "config": {"workload": "C2: ordinary point kriging, 100k samples, 1M block targets, <=32 neighbours, "
                               "nested exponential + spherical with per-structure ranges",
                   "targets": m, "samples": w.n, "max_samples": 32, "parallelism": "single GPU"},
        "fp64": rep["fp64"],
        "speedup_fp32ir_over_fp64": rep["speedup_fp32ir_over_fp64"],
        "max_rel_err_vs_fp64": {"value": rep["rel_err_fp32ir_vs_fp64"]["value"]["max"],
                                "variance": rep["rel_err_fp32ir_vs_fp64"]["variance"]["max"],
                                "floor": rep["rel_err_floor"]},
        "status_equal": rep["status_equal_fp32ir_fp64"],
        "stages_ms_fp32ir": rep["stages_ms_fp32ir"],
        "stages_ms_fp64": rep["stages_ms_fp64"],
        "roofline": roof,
        "roofline_other": other,
        "cpu_baseline": rep.get("cpu_baseline"),
        "loo_cv_grad": configs.get("C4"),
        "mik": mik,
        "sgs": sgs,
        "configs": configs,
        "graph_launches": graph_launches,
    }
    print(json.dumps(line), flush=True)
    if args.copy_mib > 0:
        src = torch.empty(args.copy_mib * 1024 * 1024 // 8, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()


if __name__ == "__main__":
    main()
