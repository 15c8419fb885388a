"""Benchmark of the recursive triangular solve on one MI355X.

Workload (the drop-in case): FP64 left-side lower non-transposed TRSM through
unified_rectrxm!, A 8192 x 8192 lower triangular (off-diagonals U(-1,1)/sqrt(n),
diagonal 2+U(0,1)), B 8192 x 8192 N(0,1), alpha = 1.5, inputs generated in HBM.
A step restores B from a resident copy and solves in place; `value` is
n^2*m flops over the mean solve time (HIP events on the solve's stream).

Also reported on the same line:
  roofline      the dominant kernel (the MFMA GEMM update): algorithmic flops
                over its own time, from per-launch HIP events, vs FP64 peak;
  cpu_baseline  the oracle (C restatement of the reference algorithm, OpenMP
                over the box's host cores) on a bounded sample: the same A with
                a subset of the right-hand sides;
  extra         FP32 and TRMM rates, the whole-operation roofline fraction and
                the backward error of the timed solve.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "nextla.jl_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

PEAK = {"f64": 78.6, "f32": 157.3}  # dense TFLOP/s, MI355X (vector = matrix for FP64)
HBM_TBS = 8.0
METRIC = "TRSM TFLOP/s + %roofline, refined-mixed vs plain FP64 at 1e-14 bwd err; p99 lat"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def solve_rate(nla, torch, n, m, side, uplo, trans, func, alpha, dtype, steps, warmup, seed=1):
    """Mean time (ms) of one in-place call over `steps` timed calls; B is
    restored from a resident copy before each call (not inside the events)."""
    from nextla_amd import inputs
    A, B0 = inputs.device_workload1(n, m, uplo, side, dtype, seed)
    B = B0.clone()
    s = torch.cuda.current_stream()
    for _ in range(warmup):
        B.copy_(B0)
        nla.unified_rectrxm_(side, uplo, trans, alpha, func, A, B)
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    t0 = time.perf_counter()
    for e0, e1 in evs:
        B.copy_(B0)
        e0.record(s)
        nla.unified_rectrxm_(side, uplo, trans, alpha, func, A, B)
        e1.record(s)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) * 1e3 / steps
    times = sorted(e0.elapsed_time(e1) for e0, e1 in evs)
    return A, B0, B, times, wall


def backward_error(torch, side, uplo, trans, alpha, A, B0, X):
    T = torch.tril(A.double()) if uplo == "L" else torch.triu(A.double())
    if trans != "N":
        T = T.t()
    Xd, Bd = X.double(), B0.double()
    R = alpha * Bd - (T @ Xd if side == "L" else Xd @ T)
    den = torch.linalg.norm(T) * torch.linalg.norm(Xd) + abs(alpha) * torch.linalg.norm(Bd)
    return (torch.linalg.norm(R) / den).item()


def cpu_baseline(A_host, n, seconds_target=12.0):
    """The oracle on host cores over a bounded number of RHS columns: a short
    calibration run at 128 columns, then a sample sized for ~seconds_target."""
    import numpy as np
    from oracle import oracle as O
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    threads = max(1, min(16, threads))
    rng = np.random.default_rng(3)
    m_try = 128
    B = rng.standard_normal((n, m_try))
    t0 = time.perf_counter()
    O.rectrxm("L", "L", "N", 1.5, "S", A_host, B, nthreads=threads)
    dt = time.perf_counter() - t0
    m_s = int(max(128, min(8192, m_try * seconds_target / max(dt, 1e-3))))
    m_s = (m_s // 32) * 32
    B = rng.standard_normal((n, m_s))
    t0 = time.perf_counter()
    X = O.rectrxm("L", "L", "N", 1.5, "S", A_host, B, nthreads=threads)
    dt = time.perf_counter() - t0
    out = {"value": n * n * m_s / dt / 1e12, "unit": "TFLOP/s", "cores": threads, "kind": "port",
           "sample": f"oracle restatement of unified_rectrxm! (thresholds 256/16, 32x32 GEMM tiles, OpenMP over "
                     f"work-groups), FP64 LL-N TRSM on the same A (n={n}) with {m_s} of the 8192 RHS columns, "
                     f"{dt:.1f} s"}
    # the reference tests' oracle library on the same sample, for context
    try:
        t0 = time.perf_counter()
        Y = O.blas("L", "L", "N", 1.5, "S", A_host, B)
        dtb = time.perf_counter() - t0
        out["openblas_tflops"] = n * n * m_s / dtb / 1e12
        out["openblas_vs_port_rel_diff"] = O.rel_error(X, Y)
    except Exception as e:  # scipy missing on a box: report the port only
        out["openblas_error"] = str(e)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=8192)
    ap.add_argument("--m", type=int, default=8192)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-extra", action="store_true", help="skip FP32/TRMM extra rates")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the path is single-GPU by design "
              f"(sharding plan in DESIGN.md), so no multi-GPU figure is produced.", flush=True)
        sys.exit(2)

    import torch
    import nextla_amd as nla
    from nextla_amd import _lib

    nla.load()
    n, m = args.n, args.m
    flops = float(n) * n * m
    log(f"bench: FP64 LL-N TRSM n={n} m={m} steps={args.steps} warmup={args.warmup}")
    A, B0, X, times, wall = solve_rate(nla, torch, n, m, "L", "L", "N", "S", 1.5, torch.float64, args.steps,
                                       args.warmup)
    mean_ms = sum(times) / len(times)
    value = flops / (mean_ms * 1e-3) / 1e12
    p99 = times[min(len(times) - 1, int(math.ceil(0.99 * len(times))) - 1)]
    bwd = backward_error(torch, "L", "L", "N", 1.5, A, B0, X)
    log(f"bench: {value:.2f} TFLOP/s, mean {mean_ms:.3f} ms, bwd err {bwd:.2e}")

    # per-kernel timing pass for the dominant kernel (GEMM update)
    _lib.prof_enable(True)
    B = B0.clone()
    for _ in range(3):
        B.copy_(B0)
        nla.unified_rectrxm_("L", "L", "N", 1.5, "S", A, B)
    torch.cuda.synchronize()
    g_n, g_ms, g_fl, g_by = _lib.prof_summary(_lib.NLA_KIND_GEMM)
    l_n, l_ms, l_fl, l_by = _lib.prof_summary(_lib.NLA_KIND_LEAF)
    _lib.prof_enable(False)
    g_tf = g_fl / (g_ms * 1e-3) / 1e12 if g_ms > 0 else 0.0
    roofline = {"bound": "mfma", "achieved": round(g_tf, 3), "peak": PEAK["f64"], "unit": "TFLOP/s",
                "frac": round(g_tf / PEAK["f64"], 4), "traffic": None,
                "kernel": "gemm_kernel<double> (MFMA f64 16x16x4 GEMM update)",
                "launches_per_solve": g_n // 3, "avg_launch_ms": round(g_ms / max(g_n, 1), 4),
                "gemm_time_share": round(g_ms / max(g_ms + l_ms, 1e-9), 4)}
    leaf = {"launches_per_solve": l_n // 3, "avg_launch_ms": round(l_ms / max(l_n, 1), 4),
            "gflops": round(l_fl / (l_ms * 1e-3) / 1e9, 1) if l_ms > 0 else None}
    q_bytes = (n * (n + 1) / 2 + 2.0 * n * m) * 8
    t_roof = max(flops / (PEAK["f64"] * 1e12), q_bytes / (HBM_TBS * 1e12))
    extra = {"op_roofline_frac": round(t_roof / (mean_ms * 1e-3), 4), "median_ms": times[len(times) // 2],
             "p99_ms": p99, "wall_ms_per_step_incl_restore": round(wall, 3), "backward_error": bwd,
             "leaf": leaf}
    del X, B
    if not args.no_extra:
        for label, dt, func in (("f32_trsm", torch.float32, "S"), ("f64_trmm", torch.float64, "M"),
                                ("f32_trmm", torch.float32, "M")):
            _, _, _, t2, _ = solve_rate(nla, torch, n, m, "L", "L", "N", func, 1.5, dt, max(3, args.steps // 2), 2)
            mm = sum(t2) / len(t2)
            extra[label + "_tflops"] = round(flops / (mm * 1e-3) / 1e12, 3)
            log(f"bench: {label} {extra[label + '_tflops']} TFLOP/s")
    torch.cuda.synchronize()

    cpu = None
    if not args.no_cpu:
        A_host = A.cpu().numpy()
        log("bench: CPU baseline (oracle) ...")
        cpu = cpu_baseline(A_host, n)
        log(f"bench: CPU {cpu['value']:.4f} TFLOP/s on {cpu['cores']} cores")

    line = {"metric": METRIC, "value": round(value, 3), "unit": "TFLOP/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(mean_ms, 4), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded, generated in HBM)",
            "config": {"workload": "plain FP64 TRSM drop-in, left lower no-transpose, alpha 1.5", "n": n, "m": m,
                       "side": "L", "uplo": "L", "trans": "N", "func": "S", "parallelism": "single GPU"},
            "roofline": roofline, "cpu_baseline": cpu, "extra": extra}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
