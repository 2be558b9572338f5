#!/usr/bin/env python
"""Benchmark of the H-matrix hot path on one MI355X.

Workload (one "step"): Schulz inversion X_{k+1} = (2I - X_k A) X_k from
X_0 = alpha A^T of the 2D 256^2 operator -div(a grad u) + u with lognormal
coefficient (log a Gaussian random field, sigma = 1, correlation length 0.1,
seed 0), HMat2d numLevels = 6, maxRank = 16, strong admissibility, run until
the estimate of ||I - X A||_2 drops below 1e-8 or stops decreasing (the
rank-16 truncation floor). The timed form is the mixed one (FP32 iterations
until the estimate is below 1e-3, FP64 for the rest); the plain FP64 form is
timed once beside it. Also reported: H*X on 1 and 64 FP32 right-hand sides
(HBM GB/s, roofline of the accumulation launch), one steady-state H*H product
(TFLOP/s), and the FP64 CPU oracle on the host cores.

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Schulz inverse time to ||I-AX||<1e-8 (s), mixed vs FP64; H-matvec HBM GB/s"
HBM_PEAK_GBS = 8000.0
PEAK_TFLOPS = {"f64": 78.6, "f32": 157.3}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def residual_power_estimate(A, X, n, steps=12, seed=3):
    """max_v ||(I - A X) v|| / ||v|| over a short power iteration (GPU matvecs)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(n)
    v /= np.linalg.norm(v)
    est = 0.0
    for _ in range(steps):
        w = v - A.Multiply(1.0, X.Multiply(1.0, v))
        nw = np.linalg.norm(w)
        est = max(est, nw)
        if nw == 0:
            break
        v = w / nw
    return est


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summary (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, factors from a streaming
    kernel of known size), or None if absent."""
    path = os.path.join(ROOT, "profiles", "matvec_256sq_lognormal_fp32_k1_traffic.md")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch %s:" % kernel):
                    return float(line.split(":")[1])
    except OSError:
        pass
    return None


def cpu_baseline(cfg, t_gpu_iters):
    """FP64 oracle (the reference algorithm restated in C++) on one host core,
    on a bounded sample of the same workload: the first Schulz product
    Z = -X_0 A (X_0 = alpha A^T) and one power step H*x. Later iterations cost
    more (their operands carry rank-16 blocks everywhere, X_0's are rank 0),
    so iterations x (2 first products + 48 power steps) is a LOWER bound on the
    oracle's Schulz time."""
    import numpy as np
    import oracle as orc
    args = (cfg["rowptr"], cfg["colind"], cfg["vals"])
    orc.rng_initialize()
    A = orc.HMat.from_csr(2, cfg["sizes"], cfg["num_levels"], cfg["max_rank"], True, *args)
    x = np.random.default_rng(0).standard_normal(A.n)
    t = time.perf_counter()
    for _ in range(5):
        A.multiply(x)
    t_mv = (time.perf_counter() - t) / 5
    X0 = A.transpose()
    X0.scale(1e-7)
    log("bench: cpu oracle product X_0 A ...")
    t = time.perf_counter()
    X0.matmul(A, alpha=-1.0)
    t_hh = time.perf_counter() - t
    per_iter = 2 * t_hh + 48 * t_mv
    return {"value": per_iter * t_gpu_iters, "unit": "s", "cores": 1, "kind": "port",
            "sample": "oracle FP64 on 1 core: first Schulz product -X_0 A on the 256^2 L=6 operator (%.1f s) and "
                      "one H*x (%.3f s); value = %d iterations x (2 products + 48 power steps), a lower bound "
                      "(later products carry rank-16 operands)" % (t_hh, t_mv, t_gpu_iters),
            "hh_seconds": t_hh, "matvec_seconds": t_mv, "lower_bound": True}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--stop-tol", type=float, default=1e-8)
    ap.add_argument("--max-iter", type=int, default=80)
    ap.add_argument("--switch-tol", type=float, default=1e-3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--mode", choices=["full", "matvec", "hh"], default="full",
                    help="matvec / hh: only the kernel loop (for rocprofv3 runs)")
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--config", choices=["c2", "c4"], default="c2",
                    help="c2: 2D 256^2 (the headline workload); c4: 2D 512^2, sigma=1.5, l=0.05, "
                         "numLevels 7, maxRank 32 (kernel modes only)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU; there is no multi-GPU path in this build (requested --gpus %d)" % a.gpus,
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import dmhm_amd
    from dmhm_amd import workloads as W

    dmhm_amd.init(0)
    if a.config == "c4" and a.mode == "full":
        print("the 512^2 workload is measured in the kernel modes (--mode matvec / hh)", file=sys.stderr)
        sys.exit(2)
    log("bench: building the %s operator" % ("256^2 lognormal" if a.config == "c2" else "512^2 lognormal"))
    cfg = W.config_operator(a.config, seed=0)
    args = (cfg["rowptr"], cfg["colind"], cfg["vals"])
    L, r, n = cfg["num_levels"], cfg["max_rank"], int(np.prod(cfg["sizes"]))
    A = dmhm_amd.HMat2d.from_csr(cfg["sizes"], L, r, True, *args)
    sopts = dict(stopTol=a.stop_tol, maxIterations=a.max_iter, switchTol=a.switch_tol, stall_stop=True)

    if a.mode == "matvec":
        X = A.Copy()
        X.SetToRandom(seed=1)
        # The FP64 -> FP32 conversion reads 8 B and writes 4 B per element: the
        # known-size launch that scales the FETCH_SIZE / WRITE_SIZE counters.
        X32 = X.Convert(dmhm_amd.DMHM_F32)
        ms, by = X32.profile_matvec(1, a.reps)
        print(json.dumps({"mode": "matvec", "config": a.config, "ms": ms.tolist(), "bytes": by.tolist(),
                          "GBps": float(by.sum() / (ms[0] * 1e-3) / 1e9),
                          "convert_elements": int(X.info().storageBytes // 8)}))
        return
    if a.mode == "hh":
        X = A.Copy()
        X.SetToRandom(seed=1)
        Y = A.Copy()
        Y.SetToRandom(seed=2)
        for _ in range(a.reps):
            ms, fl = X.profile_hh(Y)
        out = {"mode": "hh", "config": a.config, "ms": ms.tolist(), "flops": fl.tolist(),
               "phases": X.last_hh_phases(cap=4096)}
        if a.config == "c4":
            X32, Y32 = X.Convert(dmhm_amd.DMHM_F32), Y.Convert(dmhm_amd.DMHM_F32)
            ms32, fl32 = X32.profile_hh(Y32)
            out["f32_ms"] = ms32.tolist()
            out["f32_flops"] = fl32.tolist()
        print(json.dumps(out))
        return

    # Warm-up: plan building and workspace. Two iterations in each precision
    # reach every plan of the loop (first product with the imported operator,
    # then products of full-rank iterates).
    for w in range(a.warmup):
        log("bench: warmup %d" % w)
        X = A.Copy()
        X.SchulzInvert(numIterations=2)
        X = A.Copy()
        X.SchulzInvert(numIterations=2, mixed=True)
    dmhm_amd.load().dmhm_synchronize()

    log("bench: timing %d mixed Schulz inversion(s)" % a.steps)
    t0 = time.perf_counter()
    reps = []
    for s in range(a.steps):
        X = A.Copy()
        reps.append(X.SchulzInvert(mixed=True, **sopts))
    dmhm_amd.load().dmhm_synchronize()
    t1 = time.perf_counter()
    sec_mixed = (t1 - t0) / a.steps
    rep = reps[-1]
    res_mixed = residual_power_estimate(A, X, n)
    log("bench: mixed %.2f s, %d iterations (%d FP32), estimate %.3e, power check %.3e"
        % (sec_mixed, rep.iterations, rep.iterationsLowPrecision, rep.residualEstimate, res_mixed))

    sec64 = None
    rep64 = None
    res64 = None
    if not a.no_fp64:
        log("bench: timing the FP64 Schulz inversion")
        X64 = A.Copy()
        t = time.perf_counter()
        rep64 = X64.SchulzInvert(mixed=False, **sopts)
        dmhm_amd.load().dmhm_synchronize()
        sec64 = time.perf_counter() - t
        res64 = residual_power_estimate(A, X64, n)
        log("bench: fp64 %.2f s, %d iterations, estimate %.3e, power check %.3e"
            % (sec64, rep64.iterations, rep64.residualEstimate, res64))
        v = np.random.default_rng(9).standard_normal(n)
        diff = np.linalg.norm(X.Multiply(1.0, v) - X64.Multiply(1.0, v)) / np.linalg.norm(X64.Multiply(1.0, v))
    else:
        diff = None

    # H*X throughput on the inverse (ranks as left by the inversion).
    X32 = X.Convert(dmhm_amd.DMHM_F32)
    ms1, by1 = X32.profile_matvec(1, a.reps)
    ms64, by64 = X32.profile_matvec(64, max(3, a.reps // 4))
    msd, byd = X.profile_matvec(1, a.reps)
    gbs1 = (by1.sum()) / (ms1[0] * 1e-3) / 1e9
    acc_gbs = by1[1] / (ms1[2] * 1e-3) / 1e9
    gbs64 = (by64.sum()) / (ms64[0] * 1e-3) / 1e9
    _, fl64 = X32.matvec_bytes(64)
    tflops64 = fl64 / (ms64[0] * 1e-3) / 1e12
    gbsd = (byd.sum()) / (msd[0] * 1e-3) / 1e9
    log("bench: matvec f32 k=1 %.3f ms %.0f GB/s (accumulation launch %.0f GB/s); k=64 %.3f ms %.1f TFLOP/s; "
        "f64 k=1 %.0f GB/s" % (ms1[0], gbs1, acc_gbs, ms64[0], tflops64, gbsd))

    # One steady-state H*H product in each precision.
    hh = {}
    for name, M in (("f64", X), ("f32", X32)):
        msh, flh = M.profile_hh(M)
        msh, flh = M.profile_hh(M)
        phases = M.last_hh_phases()
        hh[name + "_phases"] = [{"kind": ["tile", "qr", "svd", "rng"][k] if k < 4 else str(k), "ms": t,
                                 "tflops": (f / (t * 1e-3) / 1e12) if (t > 0 and f > 0) else None}
                                for k, t, f in phases]
        hh[name] = {"ms": msh[0], "tile_ms": msh[1], "qr_ms": msh[2], "svd_ms": msh[3],
                    "tile_tflops": flh[0] / (msh[1] * 1e-3) / 1e12,
                    "longest_tile_launch_ms": msh[5],
                    "longest_tile_launch_tflops": flh[1] / (msh[5] * 1e-3) / 1e12 if msh[5] > 0 else 0.0}
    log("bench: H*H f64 %.1f ms, f32 %.1f ms" % (hh["f64"]["ms"], hh["f32"]["ms"]))

    cpu = None
    if not a.no_cpu:
        log("bench: CPU oracle sample (one core)")
        os.environ.setdefault("OMP_NUM_THREADS", "1")
        cpu = cpu_baseline(cfg, rep64.iterations if rep64 is not None else rep.iterations)
        log("bench: cpu %.1f s extrapolated" % cpu["value"])

    out = {
        "metric": METRIC,
        "value": sec_mixed,
        "unit": "s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": sec_mixed * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": "2D 256^2 -div(a grad u)+u, log a GRF sigma=1 l=0.1 (seed 0), HMat2d numLevels=6 "
                               "maxRank=16 strong; Schulz to ||I-XA||<1e-8 or truncation floor, FP32 iterations "
                               "then FP64 (switch at 1e-3)",
                   "n": n, "numLevels": L, "maxRank": r, "oversample": 4, "stopTol": a.stop_tol},
        "schulz": {"mixed_seconds": sec_mixed, "mixed_iterations": rep.iterations,
                   "mixed_fp32_iterations": rep.iterationsLowPrecision,
                   "mixed_residual_estimate": rep.residualEstimate, "mixed_residual_power_check": res_mixed,
                   "fp64_seconds": sec64, "fp64_iterations": rep64.iterations if rep64 else None,
                   "fp64_residual_estimate": rep64.residualEstimate if rep64 else None,
                   "fp64_residual_power_check": res64,
                   "speedup_mixed_vs_fp64": (sec64 / sec_mixed) if sec64 else None,
                   "mixed_vs_fp64_relative_difference": diff,
                   "reached_1e-8": bool(rep.residualEstimate < 1e-8)},
        "matvec": {"f32_k1_ms": ms1[0], "f32_k1_GBps": gbs1, "f32_k1_accumulation_GBps": acc_gbs,
                   "f32_k64_ms": ms64[0], "f32_k64_GBps": gbs64, "f32_k64_TFLOPs": tflops64,
                   "f64_k1_ms": msd[0], "f64_k1_GBps": gbsd, "f32_k1_bytes": by1.sum()},
        "hh_product": hh,
        "roofline": {"bound": "hbm", "kernel": "mv_acc_kernel<float,1> (H*x accumulation launch, FP32, k=1)",
                     "achieved": acc_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": acc_gbs / HBM_PEAK_GBS,
                     "traffic": measured_traffic("mv_acc_kernel<float,1>"), "algorithmic_bytes": by1[1],
                     "launch_ms": ms1[2]},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
