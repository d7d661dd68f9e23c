#!/usr/bin/env python3
"""Benchmark of the radial multipole integrals I_L on one MI355X.

Workload "kernel_only" (default for now): one channel pair, exponential grid
r_j = 1e-5 exp(j h), N = 15000, r_max = 1000 a0, Simpson weights,
rho_a = u_1s u_2p, rho_b = chi_f chi_i (seeded k, delta, l), L = 0..30,
direct + exchange (62 integrals per step), plain FP64 contraction.

A step is one pass of the hot path over the batch with inputs resident in
HBM.  Prints ONE JSON line (see the `result` dict at the end).
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "I_L integrals/sec (all L, dir+exch) at N=15000, FP32-MFMA vs FP64; rel err"
PEAK_FP64 = 78.6e12      # MI355X FP64 (vector and matrix) FLOP/s
PEAK_FP32_MFMA = 157.3e12


def cpu_info():
    model = platform.processor() or ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    blas_threads = None
    try:
        from threadpoolctl import threadpool_info
        info = threadpool_info()
        blas_threads = max((i.get("num_threads", 1) for i in info if i.get("user_api") == "blas"), default=None)
    except Exception:
        pass
    return model, os.cpu_count(), blas_threads


def plain_work_flops(n_rows, n_cols, n_L, families=2):
    """W = 2 N1 N2 min(D1,D2) + 2 N1 P per (L, family) with D1 = D2 = P = 1."""
    return families * n_L * (2.0 * n_rows * n_cols + 2.0 * n_rows)


def bench_kernel_only(args):
    import torch
    from dwba_amd import _hip, synthetic as S

    N, L_max = args.n, 30
    g, (ad, bd, ae, be, draws) = S.kernel_only_workload(N, seed=args.seed)
    dev = torch.device("cuda:0")
    t = lambda x: torch.as_tensor(np.ascontiguousarray(x), dtype=torch.float64, device=dev)
    r_d, ad_d, bd_d, ae_d, be_d = map(t, (g.r, ad, bd, ae, be))
    _hip.ensure_init(0)
    lib = _hip.load_library()
    ws_bytes = lib.dwba_il_plain_f64_workspace_bytes(N, N, L_max)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
    out = torch.empty((2, L_max + 1), dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream()
    sh = int(stream.cuda_stream)
    P = _hip.ptr

    def step():
        _hip.check(lib.dwba_il_plain_f64(P(r_d), N, N, P(ad_d), P(bd_d), P(ae_d), P(be_d), L_max, P(out),
                                         P(ws), ws_bytes, sh), "dwba_il_plain_f64")

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t_wall = time.perf_counter()
    e0.record(stream)
    for _ in range(args.steps):
        step()
    e1.record(stream)
    torch.cuda.synchronize()
    t_wall = time.perf_counter() - t_wall
    ms = e0.elapsed_time(e1) / args.steps
    n_int = 2 * (L_max + 1)
    value = n_int / (ms * 1e-3)
    W = plain_work_flops(N, N, L_max + 1)
    res = out.cpu().numpy()

    # parity against the oracle's exact O(N) long-double contraction
    from oracle import semisep
    ref_d = semisep.il_all_L(g.r, ad, bd, L_max)
    ref_e = semisep.il_all_L(g.r, ae, be, L_max)
    rel = max(float(np.max(np.abs(res[0] - ref_d) / np.abs(ref_d))),
              float(np.max(np.abs(res[1] - ref_e) / np.abs(ref_e))))

    result = {
        "metric": METRIC, "value": value, "unit": "integrals/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "kernel-only single channel pair, exp grid N=15000 on [1e-5,1000] a0, "
                               "L=0..30, direct+exchange, plain FP64 (Simpson weights)",
                   "N": N, "L_max": L_max, "integrals_per_step": n_int, "seed": args.seed, "draws": draws},
        "roofline": {"bound": "mfma", "achieved": W / (ms * 1e-3) / 1e12, "peak": PEAK_FP64 / 1e12,
                     "unit": "TFLOP/s", "frac": W / (ms * 1e-3) / PEAK_FP64, "traffic": None,
                     "kernel": "plain_f64_kernel", "work_flop_per_launch": W,
                     "note": "FP64 VALU path; MI355X FP64 vector peak = FP64 matrix peak = 78.6 TF"},
        "rel_err_vs_fp64_oracle": rel,
        "wall_s_timed": t_wall,
    }
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline_kernel_only(g, ad, bd, ae, be, L_max)
    return result


def cpu_baseline_kernel_only(g, ad, bd, ae, be, L_max, sample_L=(0, 1, 2, 3)):
    """Oracle's faithful restatement of the reference's plain branch (full
    N x N ratio/inv_gtr/log_ratio, per-L N^2 exp + GEMVs), timed on a sample
    of multipoles and extrapolated to L = 0..L_max."""
    from oracle import radial
    model, ncpu, blas = cpu_info()
    t0 = time.perf_counter()
    radial.plain_faithful(g.r, ad, bd, ae, be, [])            # setup only
    t_setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    radial.plain_faithful(g.r, ad, bd, ae, be, list(sample_L))
    t_sample = time.perf_counter() - t0
    per_L = max(t_sample - t_setup, 1e-9) / len(sample_L)
    t_full = t_setup + per_L * (L_max + 1)
    return {"value": 2 * (L_max + 1) / t_full, "unit": "integrals/s",
            "cores": int(blas or 1), "kind": "port",
            "sample": f"one call's N x N setup ({t_setup:.1f} s) + L={list(sample_L)} "
                      f"({per_L:.2f} s per L) at N={len(g.r)}, extrapolated to L=0..{L_max}; "
                      f"numpy exp single-threaded, BLAS GEMV threads={blas}; host {model}, "
                      f"os.cpu_count={ncpu}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=15000)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--workload", default="kernel_only", choices=["kernel_only"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the I_L path is built and timed on a single MI355X; "
              "multi-GPU sharding is described in DESIGN.md, not run.", file=sys.stderr)
        sys.exit(2)
    result = bench_kernel_only(args)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
