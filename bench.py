#!/usr/bin/env python
"""Kriged targets per second of the MI355X engine on the 1M-target workload.

Workload (synthetic, seeded): 100,000 drill-hole composites (400 holes, 2 m
composites to -500 m), nested variogram nugget 0.1 + exponential (0.5;
150/100/40 m) + spherical (0.4; 400/250/80 m) rotated (30, 20, 0), ellipsoid
search 400/250/80 m with at most 32 neighbours, 200 x 200 x 25 block
centroids (1,000,000 targets). One step = one pass of the hot path over all
targets with inputs resident in HBM: device grid build, neighbour search,
bucketing, batched solves (and the deferred systems). The headline `value`
is the FP32 + FP64-refinement form; the plain FP64 form runs beside it for the
speed-up and the error of the refined form against it.

Prints one JSON line. `roofline` is for the kernel with the most time in the
headline form, timed live with HIP events on the library's stream;
`cpu_baseline` times the oracle (the C++ restatement of the reference's CPU
path, the reference itself cannot be built here) on a bounded sample of the
same targets with every host thread it may use.
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
PEAK_FP32_TFLOPS = 157.3  # MI355X FP32
PEAK_HBM_GBS = 8000.0

# Per-structure covariance cost in flops (a transcendental counts as one).
SHAPE_COST = {0: 8, 1: 6, 2: 7, 3: 14, 4: 11, 5: 12, 6: 6, 7: 8, 8: 6}


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


def search_bytes_per_target(n):
    """Least HBM traffic of one search: the target (24 B), the selected
    samples' projected coordinates and index (32 B each), the list written (4 B each)."""
    return 24.0 + 36.0 * np.asarray(n, dtype=np.float64)


def measured_traffic(key):
    """HBM bytes per launch recorded by the PMC summary in profiles/ (the
    FETCH_SIZE and WRITE_SIZE passes scaled by the copy calibration), or None."""
    path = os.path.join(ROOT, "profiles", "k_krige32_c2_pmc.md")
    try:
        with open(path) as f:
            for ln in f:
                if ln.startswith(key + ":"):
                    return float(ln.split(":", 1)[1])
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--targets", type=int, default=0, help="use the first N targets only (0: all)")
    ap.add_argument("--loo-steps", type=int, default=200, help="Adam steps of the LOO-CV gradient leg (0: skip)")
    ap.add_argument("--no-autotune", action="store_true", help="keep the library's default tuning")
    ap.add_argument("--mik-steps", type=int, default=3,
                    help="timed passes of the multiple indicator kriging leg (0: skip)")
    ap.add_argument("--config-steps", type=int, default=2,
                    help="timed passes of the uneven (4M targets, octants) and block (2M blocks) legs (0: skip)")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="finish with a device-to-device copy of this many MiB (byte calibration of PMC runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the engine shards targets across GPUs "
              "by design but no multi-GPU figure is produced", file=sys.stderr)
        sys.exit(2)

    import torch

    from boitata_amd import _abi, datasets
    from boitata_amd.engine import Engine

    if not torch.cuda.is_available():
        print("no GPU", file=sys.stderr)
        sys.exit(1)
    dev = torch.device("cuda:0")
    w = datasets.config2()
    T = w.targets if args.targets <= 0 else w.targets[: args.targets]
    m = len(T)
    print(f"[bench] workload: {w.n} samples, {m} targets", file=sys.stderr, flush=True)

    tt = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    sx, sy, sz, sv = tt(w.x), tt(w.y), tt(w.z), tt(w.value)
    hole = tt(w.hole.astype(np.int32))
    tx, ty, tz = tt(T[:, 0]), tt(T[:, 1]), tt(T[:, 2])
    e = Engine(0)
    e.set_samples(sx, sy, sz, sv, None, hole)
    vg, search = w.variogram, w.search
    # Each form writes into its own output arrays, so a repeated pass has the
    # same arguments and the library runs it as a captured HIP graph.
    outs = {}

    def run(prec):
        outs[prec] = e.krige(vg, search, tx, ty, tz, precision=prec, out=outs.get(prec))
        return outs[prec]

    # The autotuner picks grid, search mode, workgroup size and bucket edges on
    # a strided sample of 64k targets (results do not depend on them).
    tuned = None
    if not args.no_autotune:
        t0 = time.perf_counter()
        tuned = e.autotune(vg, search, tx, ty, tz, precision=_abi.FP32_IR, sample=65536)
        tuned["seconds"] = time.perf_counter() - t0
        print(f"[bench] autotune: {tuned}", file=sys.stderr, flush=True)
    for _ in range(max(args.warmup, 1)):
        run(_abi.FP32_IR)
        run(_abi.FP64)
    torch.cuda.synchronize()

    def timed(prec):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = None
        for _ in range(args.steps):
            out = run(prec)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / args.steps, out

    dt32, out32 = timed(_abi.FP32_IR)
    dt64, out64 = timed(_abi.FP64)
    print(f"[bench] FP32+IR {dt32 * 1e3:.2f} ms/step, FP64 {dt64 * 1e3:.2f} ms/step", file=sys.stderr, flush=True)

    # Stage times from HIP events on the library's stream.
    e.profile(True)
    run(_abi.FP32_IR)
    st32, _ = e.profile_read()
    run(_abi.FP64)
    st64, _ = e.profile_read()
    e.profile(False)

    # Error of the refined form against plain FP64, floor 1e-3 * sill.
    c0 = vg.nugget + sum(vg.structures[k].sill for k in range(vg.n_structures))
    s32, s64 = out32["status"].cpu().numpy(), out64["status"].cpu().numpy()
    ok = (s32 == 0) & (s64 == 0)
    v32, v64 = out32["value"].cpu().numpy()[ok], out64["value"].cpu().numpy()[ok]
    q32, q64 = out32["variance"].cpu().numpy()[ok], out64["variance"].cpu().numpy()[ok]
    floor = 1e-3 * c0
    err_v = float(np.max(np.abs(v32 - v64) / np.maximum(np.abs(v64), floor))) if ok.any() else float("nan")
    err_q = float(np.max(np.abs(q32 - q64) / np.maximum(np.abs(q64), floor))) if ok.any() else float("nan")
    status_equal = bool((s32 == s64).all())

    n_used = out64["n_used"].cpu().numpy()
    shape_sum = sum(SHAPE_COST[vg.structures[k].model] for k in range(vg.n_structures))
    flops = float(flops_per_system(n_used[s64 == 0], shape_sum).sum())
    sbytes = float(search_bytes_per_target(n_used).sum())

    # The dominant kernel of the headline form. Its HBM traffic per launch comes
    # from the committed FETCH_SIZE / WRITE_SIZE passes on this workload (only
    # quoted when the full 1M-target workload runs).
    solve_ms = st32["solve"]
    search_ms = st32["search"]
    traffic = measured_traffic("traffic_bytes_k_krige32_c2") if m == len(w.targets) else None
    if solve_ms >= search_ms:
        ach = flops / (solve_ms * 1e-3) / 1e12
        roof = dict(bound="mfma", achieved=ach, peak=PEAK_FP64_TFLOPS, unit="TFLOP/s", frac=ach / PEAK_FP64_TFLOPS,
                    traffic=traffic, kernel="k_krige32<32> (FP32 Gauss-Jordan inverse, FP64 refinement)",
                    note="compute (VALU) bound batched solves: algorithmic F(n) flops per system over the kernel's "
                         "time against the FP64 vector peak (= the FP64 matrix peak on MI355X); packed FP32 is off")
    else:
        ach = sbytes / (search_ms * 1e-3) / 1e9
        roof = dict(bound="hbm", achieved=ach, peak=PEAK_HBM_GBS, unit="GB/s", frac=ach / PEAK_HBM_GBS,
                    traffic=None, kernel="k_knn_wave")
    other = dict(search_gbs=sbytes / (search_ms * 1e-3) / 1e9 if search_ms > 0 else None,
                 solve32_tflops=flops / (solve_ms * 1e-3) / 1e12 if solve_ms > 0 else None,
                 solve64_tflops=flops / (st64["solve"] * 1e-3) / 1e12 if st64["solve"] > 0 else None)

    # Second half of the metric: LOO-CV gradient steps/s on 200,000 samples of
    # the same generator (800 holes), 12 parameters, fixed neighbour sets of
    # the explicit 400/250/80 ellipsoid (<= 32 neighbours), Adam from
    # parameters 30% off the generating values.
    from boitata_amd import loocv

    loo_line = None
    if args.loo_steps > 0:
        w4 = datasets.config4()
        models = ["exponential", "spherical"]
        e.set_samples(tt(w4.x), tt(w4.y), tt(w4.z), tt(w4.value), None, tt(w4.hole.astype(np.int32)))
        e.loo_prepare(w4.search, None)
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
        # The same steps with the FP64 forward and adjoint, and the agreement of
        # the two at the final parameters.
        dtl64, _, hist64 = adam_rate(_abi.FP64, max(args.loo_steps // 4, 1))
        loss, grad, used = loocv.loss_and_grad(e, th, models, _abi.FP32_IR)
        loss64, grad64, _ = loocv.loss_and_grad(e, th, models, _abi.FP64)
        gfloor = 1e-3 * float(np.linalg.norm(grad64))
        loo_line = {"value": args.loo_steps / dtl, "unit": "Adam steps/s", "steps": args.loo_steps,
                    "ms_per_step": dtl / args.loo_steps * 1e3, "samples": w4.n, "systems_per_step": used,
                    "parameters": len(th0), "lr": 0.02, "loss_first": hist[0], "loss_last": hist[-1],
                    "precision": "FP32 Gauss-Jordan forward with FP64 refinement, FP32 adjoint, FP64 reduction",
                    "fp64": {"value": max(args.loo_steps // 4, 1) / dtl64, "unit": "Adam steps/s"},
                    "rel_err_vs_fp64": {"loss": abs(loss - loss64) / abs(loss64),
                                        "grad": float(np.max(np.abs(grad - grad64) /
                                                             np.maximum(np.abs(grad64), gfloor)))}}
        print(f"[bench] LOO-CV: {loo_line['value']:.1f} Adam steps/s on {w4.n} samples, loss {hist[0]:.4f} -> "
              f"{hist[-1]:.4f}", file=sys.stderr, flush=True)
        e.set_samples(sx, sy, sz, sv, None, hole)

    # Multiple indicator kriging on the same samples and targets: the nine
    # deciles of the data as thresholds with the one variogram (median
    # indicator kriging), three cutoffs and three quantiles; one pass = search,
    # solves writing weights, indicator estimates, conditional distributions.
    mik_line = None
    if args.mik_steps > 0:
        thr = np.quantile(w.value, np.linspace(0.1, 0.9, 9))
        mk = _abi.make_mik(thr, 1)
        kw = dict(cutoffs=[0.5, 1.0, 2.0], quantiles=[0.1, 0.5, 0.9])
        mres = {}
        for prec in (_abi.FP32_IR, _abi.FP64):
            mres[prec] = e.mik_predict(mk, thr, [vg], search, tx, ty, tz, precision=prec, **kw)  # warm-up
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.mik_steps):
                mres[prec] = e.mik_predict(mk, thr, [vg], search, tx, ty, tz, precision=prec, **kw)
            torch.cuda.synchronize()
            mres[prec]["dt"] = (time.perf_counter() - t0) / args.mik_steps
        a32, a64 = mres[_abi.FP32_IR], mres[_abi.FP64]
        okm = ((a32["status"] == 0) & (a64["status"] == 0)).cpu().numpy()
        def rel(k):
            x, y = a32[k].cpu().numpy()[..., okm], a64[k].cpu().numpy()[..., okm]
            return float(np.max(np.abs(x - y) / np.maximum(np.abs(y), 1e-3))) if okm.any() else float("nan")
        mik_line = {"value": m / a32["dt"], "unit": "targets/s", "ms_per_step": a32["dt"] * 1e3,
                    "steps": args.mik_steps, "thresholds": len(thr), "variograms": 1, "cutoffs": 3, "quantiles": 3,
                    "precision": "FP32 Gauss-Jordan weights with FP64 refinement",
                    "fp64": {"value": m / a64["dt"], "ms_per_step": a64["dt"] * 1e3},
                    "rel_err_vs_fp64": {"cdf": rel("cdf"), "mean": rel("mean"), "variance": rel("variance")},
                    "estimated": int(okm.sum())}
        print(f"[bench] MIK: {mik_line['value']:.0f} targets/s FP32+IR, {mik_line['fp64']['value']:.0f} FP64",
              file=sys.stderr, flush=True)

    # The other workloads, measured beside the headline: uneven neighbourhoods
    # (4M targets among clustered samples, octant search, 4-64 neighbours:
    # size buckets) and block kriging (2M blocks discretised 4 x 4 x 2, 64
    # neighbours, Gaussian covariance with a 1e-4 nugget: ill-conditioned
    # systems that the refined form hands back to FP64).
    configs = None
    if args.config_steps > 0:
        configs = {}
        for name, gen in (("uneven_4m_octants", datasets.config3), ("block_2m_gaussian", datasets.config5)):
            wc = gen()
            Tc = wc.targets
            mc = len(Tc)
            holes = None if wc.hole is None else tt(wc.hole.astype(np.int32))
            ec = Engine(0)  # each workload on a context of its own (its own tuning)
            ec.set_samples(tt(wc.x), tt(wc.y), tt(wc.z), tt(wc.value), None, holes)
            cx, cy, cz = tt(Tc[:, 0]), tt(Tc[:, 1]), tt(Tc[:, 2])
            blk = getattr(wc, "block", None)
            tuned_c = None
            if not args.no_autotune and blk is None:
                tuned_c = ec.autotune(wc.variogram, wc.search, cx, cy, cz, precision=_abi.FP32_IR, sample=65536)
            res = {}
            for prec in (_abi.FP32_IR, _abi.FP64):
                def one():
                    if blk is None:
                        return ec.krige(wc.variogram, wc.search, cx, cy, cz, precision=prec, out=res.get(prec))
                    return ec.block_krige(wc.variogram, wc.search, blk, cx, cy, cz, precision=prec)
                res[prec] = one()
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                for _ in range(args.config_steps):
                    res[prec] = one()
                torch.cuda.synchronize()
                res[prec]["dt"] = (time.perf_counter() - t0) / args.config_steps
            a32, a64 = res[_abi.FP32_IR], res[_abi.FP64]
            s32, s64 = a32["status"].cpu().numpy(), a64["status"].cpu().numpy()
            okc = (s32 == 0) & (s64 == 0)
            c0c = wc.variogram.nugget + sum(wc.variogram.structures[k].sill for k in range(wc.variogram.n_structures))
            errs = {}
            for k in ("value", "variance"):
                x, y = a32[k].cpu().numpy()[okc], a64[k].cpu().numpy()[okc]
                errs[k] = float(np.max(np.abs(x - y) / np.maximum(np.abs(y), 1e-3 * c0c))) if okc.any() else None
            nu = a64["n_used"].cpu().numpy()[s64 == 0]
            configs[name] = {
                "targets": mc, "samples": wc.n, "value": mc / a32["dt"], "unit": "targets/s",
                "ms_per_step": a32["dt"] * 1e3, "fp64": {"value": mc / a64["dt"], "ms_per_step": a64["dt"] * 1e3},
                "speedup_fp32ir_over_fp64": a64["dt"] / a32["dt"], "max_rel_err_vs_fp64": errs,
                "status_equal": bool((s32 == s64).all()), "estimated": int((s64 == 0).sum()),
                "n_used": {"min": int(nu.min()) if len(nu) else 0, "max": int(nu.max()) if len(nu) else 0,
                           "le16": int((nu <= 16).sum()), "le32": int(((nu > 16) & (nu <= 32)).sum()),
                           "le64": int((nu > 32).sum())},
                "support": "block 4x4x2" if blk is not None else "point",
                "refine_limit": tuned_c["refine_limit"] if tuned_c else None}
            print(f"[bench] {name}: {configs[name]['value']:.0f} targets/s FP32+IR, "
                  f"{configs[name]['fp64']['value']:.0f} FP64", file=sys.stderr, flush=True)
            ec.close()
            del res

    cpu = None
    if not args.no_cpu:
        import oracle

        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        d = oracle.Data(w.x, w.y, w.z, w.value, None, w.hole)
        probe = T[:: max(m // 4000, 1)][:4000]
        t0 = time.perf_counter()
        oracle.krige(d, vg, search, probe, threads=threads)
        rate = len(probe) / (time.perf_counter() - t0)
        k = int(min(m, max(len(probe), rate * args.cpu_seconds)))
        sample = T[:: max(m // k, 1)][:k]
        # Whole passes over the sample until about cpu_seconds of work.
        passes, done, t0 = 0, 0, time.perf_counter()
        while True:
            oracle.krige(d, vg, search, sample, threads=threads)
            passes += 1
            done += len(sample)
            dt = time.perf_counter() - t0
            if dt >= args.cpu_seconds or passes >= 64:
                break
        cpu = dict(value=done / dt, unit="targets/s", cores=threads, kind="port",
                   sample=f"{passes} pass(es) over {len(sample)} of the {m} targets (every {max(m // k, 1)}th), "
                          f"neighbour search + FP64 ordinary kriging, {dt:.1f} s")
        # The same restatement on one thread, over a smaller strided sample.
        one = T[:: max(m // 60000, 1)][:60000]
        t0 = time.perf_counter()
        oracle.krige(d, vg, search, one, threads=1)
        dt1 = time.perf_counter() - t0
        cpu["single_thread"] = dict(value=len(one) / dt1, unit="targets/s", cores=1,
                                    sample=f"{len(one)} of the {m} targets (every {max(m // 60000, 1)}th), {dt1:.1f} s")
        print(f"[bench] CPU oracle {cpu['value']:.0f} targets/s on {threads} threads, "
              f"{cpu['single_thread']['value']:.0f} on one", file=sys.stderr, flush=True)
        if mik_line is not None:
            # The oracle's multiple indicator kriging on a strided sample.
            few = T[:: max(m // 20000, 1)][:20000]
            t0 = time.perf_counter()
            oracle.mik_predict(d, mk, thr, [vg], search, few, threads=threads, **kw)
            dtm = time.perf_counter() - t0
            mik_line["cpu_baseline"] = dict(value=len(few) / dtm, unit="targets/s", cores=threads, kind="port",
                                            sample=f"{len(few)} of the {m} targets, {dtm:.1f} s")

    line = {
        "metric": "kriged targets/s (FP32+IR vs FP64) at <=1e-6 rel err; LOO-CV grad steps/s",
        "value": m / dt32,
        "unit": "targets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt32 * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp64 refinement",
        "data": "synthetic (seeded drill-hole composites, lognormal Gaussian-field values)",
        "config": {"workload": "C2: ordinary point kriging, 100k samples, 1M block targets, <=32 neighbours, "
                               "nested exponential + spherical with per-structure ranges",
                   "targets": m, "samples": w.n, "max_samples": 32, "parallelism": "single GPU"},
        "fp64": {"value": m / dt64, "ms_per_step": dt64 * 1e3},
        "speedup_fp32ir_over_fp64": dt64 / dt32,
        "max_rel_err_vs_fp64": {"value": err_v, "variance": err_q, "floor": floor},
        "status_equal": status_equal,
        "stages_ms_fp32ir": st32,
        "stages_ms_fp64": st64,
        "roofline": roof,
        "roofline_other": other,
        "cpu_baseline": cpu,
        "loo_cv_grad": loo_line,
        "mik": mik_line,
        "other_workloads": configs,
        "autotune": tuned,
        "graph_launches": e.stat(_abi.STAT_GRAPH_LAUNCHES),
    }
    print(json.dumps(line), flush=True)
    if args.copy_mib > 0:
        src = torch.empty(args.copy_mib * 1024 * 1024 // 8, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
    e.close()


if __name__ == "__main__":
    main()
