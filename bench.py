#!/usr/bin/env python3
"""Benchmark of the radial multipole integrals I_L on one MI355X.

Default workload "batched" -- the batched precision sweep: 64 energies, each
with its own exponential grid r_j = 1e-5 exp(j h) on [1e-5, 1000] a0 with
N ~ U{4000..15000}, 400 channel pairs per energy (l_i, l_f ~ U{0..40}, per
energy k_i, k_f ~ U[0.5, 8], per wave delta ~ U[0, pi]), H 1s -> 2p bound
pair, Full-Split quadrature (Filon/Clenshaw-Curtis alpha/beta, FP64 outer
tails), L = 0..30, direct + exchange: 64*400*31*2 = 1,587,200 integrals per
step.  `value` is the throughput of the FP64 path whose contractions use the
O(N) semiseparable kernel; the FP32-MFMA, split-BF16x3 and dense FP64 paths
are timed on the same batch and their errors stated against the FP64 result
(per L, per family, relative to |I| and to sum |alpha_i| K_ij |beta_j|).

Workload "kernel_only": one channel pair, N = 15000, L = 0..30, FP64 (dense
plain kernel and semiseparable kernel).
Workload "backward": gradients of sum_L c_L I_L, N = 15000, 32 pairs, L = 0..20.
Workload "ionization": 48 ejected-energy points x 400 (l_i, l_f, l_ej) triples,
Slater bound state x 3 continuum waves, N ~ U{6000..15000}, L = 0..30.

A step is one pass of the hot path over the batch with inputs resident in
HBM (HIP-graph replay of prep, tails, contractions and assembly); host
planning and uploads happen before the timed region.  Prints ONE JSON line.
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
PEAK_HBM = 8.0e12        # MI355X HBM3E bytes/s


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

    # the same pair through the O(N) semiseparable kernel
    import ctypes
    pj = (_hip.PlainJob * 1)(_hip.PlainJob(P(r_d), P(ad_d), P(bd_d), P(ae_d), P(be_d), N, N))
    pj_d = _hip.struct_array_to_device(pj, dev)
    nr = (ctypes.c_int32 * 1)(N)
    ss_bytes = lib.dwba_il_semisep_f64_workspace_bytes(1, nr, nr, L_max, 0)
    ss_ws = torch.empty(ss_bytes, dtype=torch.uint8, device=dev)
    out_ss = torch.empty((1, 2, L_max + 1), dtype=torch.float64, device=dev)

    def step_ss():
        _hip.check(lib.dwba_il_semisep_f64(1, P(pj_d), nr, nr, L_max, 0, 0, P(out_ss), P(ss_ws), ss_bytes, sh),
                   "dwba_il_semisep_f64")

    for _ in range(args.warmup + 1):
        step_ss()
    ms_ss = _events_ms(step_ss, args.steps, stream)
    res_ss = out_ss.cpu().numpy()[0]

    # parity against the oracle's exact O(N) long-double contraction
    from oracle import semisep
    ref_d = semisep.il_all_L(g.r, ad, bd, L_max)
    ref_e = semisep.il_all_L(g.r, ae, be, L_max)
    rel = max(float(np.max(np.abs(res[0] - ref_d) / np.abs(ref_d))),
              float(np.max(np.abs(res[1] - ref_e) / np.abs(ref_e))))
    rel_ss = max(float(np.max(np.abs(res_ss[0] - ref_d) / np.abs(ref_d))),
                 float(np.max(np.abs(res_ss[1] - ref_e) / np.abs(ref_e))))

    result = {
        "metric": METRIC, "value": value, "unit": "integrals/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "kernel-only single channel pair, exp grid N=15000 on [1e-5,1000] a0, "
                               "L=0..30, direct+exchange, plain FP64 (Simpson weights)",
                   "N": N, "L_max": L_max, "integrals_per_step": n_int, "seed": args.seed, "draws": draws},
        "roofline": {"bound": "fp64_valu", "achieved": W / (ms * 1e-3) / 1e12, "peak": PEAK_FP64 / 1e12,
                     "unit": "TFLOP/s", "frac": W / (ms * 1e-3) / PEAK_FP64, "traffic": None,
                     "kernel": "plain_f64_kernel", "work_flop_per_launch": W,
                     "note": "FP64 VALU path; MI355X FP64 vector peak = FP64 matrix peak = 78.6 TF"},
        "rel_err_vs_fp64_oracle": rel,
        "wall_s_timed": t_wall,
        "semiseparable": {"value": n_int / (ms_ss * 1e-3), "unit": "integrals/s", "ms_per_step": ms_ss,
                          "rel_err_vs_fp64_oracle": rel_ss, "kernel": "semisep_kernel<8> (dwba_il_semisep_f64)",
                          "note": "O(N) per L; one pair fills few waves, so this time is launch latency"},
    }
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline_kernel_only(g, ad, bd, ae, be, L_max)
    return result


def bench_backward(args):
    """Gradient workload: N = 15000, 32 pairs, L = 0..20; d/d rho_a and d/d rho_b
    of sum_{p,L} c_L I_L(p).  value: the FP64 O(N) path (K_L x by the
    semiseparable recurrences for every L, dwba_il_backward_semisep_f64);
    the dense FP64 and FP32 Keff kernels are timed beside it."""
    import torch
    from dwba_amd import synthetic as S
    from dwba_amd.autograd import density_gradients
    N, P, L_max = args.n, 32, 20
    grid, ra, rb, c = S.backward_workload(N, P, L_max, seed=args.seed)
    dev = torch.device("cuda:0")
    t = lambda x: torch.as_tensor(np.ascontiguousarray(x), dtype=torch.float64, device=dev)
    r_d, w_d, ra_d, rb_d, c_d = map(t, (grid.r, grid.w_simpson, ra, rb, c))
    stream = torch.cuda.current_stream()
    # dense: per element and pair, Horner for Keff (L_max FMAs + 1 mul) and two products
    W_dense = float(P) * N * N * (2 * L_max + 1 + 4)
    # O(N): per point, multipole, pair and side, the forward and backward recurrence steps of the
    # GEMV (SS_FLOP) and one FMA of the combination
    W_ss = float(P) * 2 * (L_max + 1) * N * (SS_FLOP + 2)
    legs = {}
    for prec, con in (("fp64", "semiseparable"), ("fp64", "dense"), ("fp32", "dense")):
        fn = lambda: density_gradients(ra_d, rb_d, r_d, w_d, c_d, L_max, prec, con)
        for _ in range(max(args.warmup, 1)):
            fn()
        legs[(prec, con)] = _events_ms(fn, args.steps, stream)
    from oracle import gradient as OG
    sub = slice(0, 4)   # error check on 4 pairs (oracle cost)
    ref_a, _ = OG.density_gradients(grid.r, grid.w_simpson, ra[sub], rb[sub], c[sub], L_max)
    mag_a, _ = OG.density_gradients(grid.r, grid.w_simpson, ra[sub], rb[sub], c[sub], L_max, absolute=True)
    errs = {}
    for prec, con in legs:
        ga, gb = density_gradients(ra_d, rb_d, r_d, w_d, c_d, L_max, prec, con)
        errs[(prec, con)] = float(np.max(np.abs(ga[sub].cpu().numpy() - ref_a) / np.maximum(mag_a, 1e-300)))
    ms = legs[("fp64", "semiseparable")]
    return {
        "metric": "gradient of sum_L c_L I_L w.r.t. rho_a and rho_b, pairs/s", "value": P / (ms * 1e-3),
        "unit": "pairs/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "backward check: exp grid N=15000 on [1e-5,1000] a0, 32 pairs, L=0..20, "
                               "c_L ~ N(0,1); value = FP64 O(N) gradients", "N": N, "pairs": P, "L_max": L_max},
        "roofline": {"bound": "fp64_valu", "achieved": W_ss / (ms * 1e-3) / 1e12, "peak": PEAK_FP64 / 1e12,
                     "unit": "TFLOP/s", "frac": W_ss / (ms * 1e-3) / PEAK_FP64, "traffic": None,
                     "kernel": "backward_setup_kernel + semisep_gemv_scan_kernel + backward_combine_kernel",
                     "work_flop_per_launch": W_ss},
        "max_err_over_magnitude": {f"{p}_{c}": e for (p, c), e in errs.items()},
        "dense_fp64": {"ms": legs[("fp64", "dense")],
                       "frac_of_fp64_peak": W_dense / (legs[("fp64", "dense")] * 1e-3) / PEAK_FP64},
        "dense_fp32": {"ms": legs[("fp32", "dense")],
                       "frac_of_fp32_valu_peak": W_dense / (legs[("fp32", "dense")] * 1e-3) / 157.3e12},
    }


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


def _events_ms(fn, steps, stream):
    import torch
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(steps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


SS_FLOP = 9.0   # per (grid point, L, family): A and Z updates (2 x add+mul), q^L step (mul), 2 FMAs into the sum


def semisep_work_flop(b):
    """Algorithmic flops of one semiseparable contraction stage of ScanBatch b:
    SS_FLOP per (grid point, L, family) of the sequential recurrences over the
    points a family needs (exchange families of Filon / plain pairs end at N1
    when their direct family goes through the shared-beta path); with that
    path, the direct families of an energy cost one forward and one backward
    recurrence plus the combination (SS_FLOP per point and L) and a dot per
    pair (2 per row and L)."""
    from dwba_amd.planner import PREP_FILON, PREP_PLAIN, PREP_STANDARD
    nl = b.L_max + 1
    gemv = bool(getattr(b, "direct_gemv", False))
    shared = set()
    W_gemv = 0.0
    if gemv:
        for ei, st in enumerate(b.E):
            ks = [k for k in st.members if b.pairs[k][2].mode in (PREP_FILON, PREP_STANDARD)]
            if ks and len({id(b.pairs[k][5]) for k in st.members}) == 1:
                shared.update(ks)
                W_gemv += SS_FLOP * st.N * nl
                W_gemv += sum(2.0 * b.pairs[k][2].N1 * nl for k in ks)
    W_scan = 0.0
    bytes_scan = 0.0
    for k, (ei, _, plan, *_rest) in enumerate(b.pairs):
        if plan.rows <= 0:
            continue
        if k in shared:
            ex = plan.rows if plan.mode in (PREP_FILON, PREP_PLAIN) else plan.cols
            W_scan += SS_FLOP * ex * nl
            bytes_scan += 8.0 * (plan.rows + ex)                    # alpha_ex, beta_ex
        else:
            W_scan += SS_FLOP * 2 * plan.cols * nl
            bytes_scan += 2 * 8.0 * (plan.rows + plan.cols)         # alpha, beta of both families
    return {"flop": W_scan + W_gemv, "scan_flop": W_scan, "scan_bytes": bytes_scan, "gemv": gemv,
            "lanes": bool(getattr(b, "semisep_lanes", False))}


def tail_work_flop(b):
    """Algorithmic FP64 work of the outer-tail stage: 3 flop (one FMA into the
    sum, one multiply of the power chain x^-(L+1)) per swept node and multipole
    L = 1..L_max, over the Filon and Levin nodes each (pair, sign) sweeps
    (planner.tail_sweep_nodes: the reference's branch choice, segment counts
    and node up-scaling).  Returns (flop, nodes)."""
    from dwba_amd.dwba_matrix_elements import TAIL_OUTER
    nodes = sum(sum(v) for k, v in getattr(b, "tail_nodes", {}).items() if b.pairs[k][2].tail_kind == TAIL_OUTER)
    return 3.0 * nodes * b.L_max, nodes


def prep_bytes(b):
    """Algorithmic HBM bytes of the prep stage: every input read once and every
    output written once, FP64 -- per pair chi_i, chi_f on [0, N) and alpha_dir,
    alpha_ex, beta_ex on [0, N1); per energy u_i, u_f, w, V_core - U_i read and
    the shared beta_dir written on [0, N)."""
    tot = 0.0
    for (ei, _, plan, *_rest) in b.pairs:
        tot += 8.0 * (2 * plan.N + 3 * plan.N1)
    for st in b.E:
        tot += 8.0 * 5 * st.N
    return tot


def tail_entry_ms(b, steps, stream):
    """Average duration (HIP events on the launch stream) of the outer-tail
    entry point alone (dwba_outer_tail_f64: Filon, Levin and combine kernels)
    on the batch's tail jobs, moments already in place."""
    from dwba_amd import planner
    hip, lib = b.hip, b.lib
    sh = hip.stream_handle(stream)
    seg, dmax = planner.filon_limits()
    Dp, D_keep = hip.f64_array(b.levin[0])
    Xp, X_keep = hip.f64_array(b.levin[1])
    fn = lambda: hip.check(lib.dwba_outer_tail_f64(b.n_tail, hip.ptr(b.tail_d), b.L_max, seg, dmax, Dp, Xp,
                                                   hip.ptr(b.tail_ws), b.tail_ws_bytes, sh), "dwba_outer_tail_f64")
    fn()
    return _events_ms(fn, steps, stream)


def stage_times(b, reps=5):
    """Stage durations (ms) of eager one-stream launches with events, averaged."""
    import torch
    acc = {}
    ov = b.overlap
    b.overlap = None
    for _ in range(reps):
        ev = {}
        b.launch(events=ev)
        torch.cuda.synchronize()
        for k, v in b.stage_ms(ev).items():
            acc[k] = acc.get(k, 0.0) + v / reps
    b.overlap = ov
    return acc


def scan_kernel_ms(b, steps, stream):
    """Average duration (HIP events on the launch stream) of the batch's
    per-pair recurrence kernel alone -- the lane-serial kernel (table-driven
    when the batch uses tables) or the wavefront-scan kernel pair -- with
    the batch's tuned configuration, after a full launch has written alpha/beta."""
    hip, lib, Lm = b.hip, b.lib, b.L_max
    seg, lb = b.semisep_config
    jobs = b.plain_d_split if b.direct_gemv else b.plain_d
    sh = hip.stream_handle(stream)
    if b.semisep_lanes and b.lane_tables:
        lbv = lb or 8
        gd, ng, max_n, _ = b._lane_table_set(lbv)
        sj, ptrs, oidx, _ = b._lane_job_set(lbv, b.direct_gemv)
        hip.check(lib.dwba_semisep_lane_tables_f64(ng, hip.ptr(gd), max_n, Lm, lbv, sh), "tables")
        if int(b.lane_split) == 4:
            sg, pg, og, ng2 = b._lane_job_set(lbv, b.direct_gemv, group=b._stage_group(lbv))
            fn = lambda: hip.check(lib.dwba_il_semisep_lanes_stage_f64(ng2, hip.ptr(sg), hip.ptr(pg), hip.ptr(og), Lm,
                                                                       lbv, hip.ptr(b.out), sh), "lanes_stage")
            nbp = 64 // b._stage_group(lbv)
            name = f"semisep_lane_stage_kernel<{lbv}, {nbp}, {2 if lbv == 4 else 3}, {4 if lbv == 4 else 2}>"
        elif b.lane_split:
            occ = 3 if int(b.lane_split) == 3 else 2
            fn = lambda: hip.check(lib.dwba_il_semisep_lanes_split_f64(b.plain_count, hip.ptr(sj), hip.ptr(ptrs),
                                                                       hip.ptr(oidx), Lm, lbv, occ, hip.ptr(b.out),
                                                                       sh), "lanes_split")
            name = f"semisep_lane_split_kernel<{lbv},*,{occ}>"
        else:
            fn = lambda: hip.check(lib.dwba_il_semisep_lanes_tab_f64(b.plain_count, hip.ptr(sj), hip.ptr(ptrs),
                                                                     hip.ptr(oidx), Lm, lbv, hip.ptr(b.out), sh),
                                   "lanes_tab")
            name = f"semisep_lane_tab_kernel<{lbv},*>"
    elif b.semisep_lanes:
        fn = lambda: hip.check(lib.dwba_il_semisep_lanes_f64(b.plain_count, hip.ptr(jobs), Lm, lb, hip.ptr(b.out),
                                                             sh), "lanes")
        name = f"semisep_lane_kernel<{lb or 8},*>"
    else:
        nra, nca = b.plain_n_split if b.direct_gemv else b.plain_n
        fn = lambda: hip.check(lib.dwba_il_semisep_f64(b.plain_count, hip.ptr(jobs), nra, nca, Lm, seg, lb,
                                                       hip.ptr(b.out), hip.ptr(b.plain_ws), b.plain_ws_bytes, sh),
                               "semisep")
        name = "semisep_plan + semisep_kernel + semisep_join"
    fn()
    ms = _events_ms(fn, steps, stream)
    b.launch()                     # leave a complete result in b.out
    return name, ms


def _source_sha16(name):
    import hashlib
    with open(os.path.join(ROOT, "dwba_amd", "csrc", name), "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()[:16]


def _committed_traffic(kernels, workload):
    """HBM bytes per launch of a stage's kernels measured with rocprofv3 --pmc
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration) and
    recorded in profiles/traffic.json together with the SHA-256 of the source
    file each kernel was built from.  Returns (bytes summed over `kernels`,
    provenance); bytes is None when a kernel has no record or its source has
    changed since it was measured."""
    names = [kernels] if isinstance(kernels, str) else list(kernels)
    p = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(p) as f:
            recs = json.load(f).get(workload, {}).get("kernels", {})
    except (OSError, ValueError):
        recs = {}
    total, prov = 0.0, []
    for k in names:
        rec = recs.get(k)
        if not rec or "source" not in rec:
            return None, f"no counter record for {k}"
        sha = _source_sha16(rec["source"])
        if sha != rec.get("source_sha16"):
            return None, f"stale: {rec['source']} changed since the counters were taken ({rec.get('source_sha16')})"
        total += rec["fetch_bytes"] + rec["write_bytes"]
        prov.append(f"{k}: {rec['source']} sha256 {sha}")
    return total, "; ".join(prov) + " (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE; profiles/)"


def _scan_leg(scan, L_max, precision, contraction, steps, warmup, graph, stream, tune=True):
    """Plan/build one ScanBatch (autotuned when it has tunable kernels), time
    `steps` graph replays after `warmup`, and take per-stage times from one
    eager launch with events."""
    import torch
    from dwba_amd.batch import ScanBatch
    b = ScanBatch(scan, L_max, precision=precision, oscillatory_method="full_split", contraction=contraction)
    if tune:
        b.tune()
    b.launch()
    ev = {}
    b.launch(events=ev)
    torch.cuda.synchronize()
    stages = b.stage_ms(ev)
    if graph:
        b.capture()
    for _ in range(warmup):
        b.run()
    ms = _events_ms(b.run, steps, stream)
    return b, ms, stages, b.results()


def _err_stats(res, ref, mag):
    d = np.abs(res - ref)
    rel = d / np.maximum(np.abs(ref), 1e-300)
    canc = d / np.maximum(mag, 1e-300)
    return {
        "exchange_max_abs_over_sum_abs_terms": float(np.max(canc[:, 1, :])),
        "exchange_max_abs_over_sum_abs_terms_per_L": [float(x) for x in np.max(canc[:, 1, :], axis=0)],
        "direct_max_abs_over_sum_abs_terms": float(np.max(canc[:, 0, :])),
        "direct_max_abs_over_sum_abs_terms_per_L": [float(x) for x in np.max(canc[:, 0, :], axis=0)],
        "exchange_median_rel": float(np.median(rel[:, 1, :])),
        "exchange_frac_rel_le_1e-6": float(np.mean(rel[:, 1, :] <= 1e-6)),
        "direct_max_rel": float(np.max(rel[:, 0, :])),
    }


def bench_batched(args, ionization=False):
    import torch
    from dwba_amd import synthetic as S

    L_max = 30
    t0 = time.perf_counter()
    if ionization:
        scan = S.ionization_scan(n_points=args.energies, n_triples=args.pairs, n_range=(6000, 15000), l_max=30,
                                 seed=args.seed)
    else:
        scan = S.excitation_scan(n_energies=args.energies, n_pairs=args.pairs, n_range=(4000, 15000), l_max=40,
                                 k_range=(0.5, 8.0), seed=args.seed)
    t_gen = time.perf_counter() - t0
    n_pairs = sum(len(e.pairs) for e in scan)
    n_int = n_pairs * (L_max + 1) * 2
    log = lambda m: print(f"[bench] {m}", file=sys.stderr, flush=True)
    # the CPU baseline's worker pool forks: run it before anything touches the GPU
    cpu = None if args.no_cpu_baseline else cpu_baseline_batched(scan, L_max)
    if cpu:
        log(f"cpu baseline: {cpu['value']:.3g} integrals/s ({cpu['sample']})")
    stream = torch.cuda.current_stream()
    graph = not args.no_graph
    log(f"scan generated: {len(scan)} energies, {n_pairs} pairs ({t_gen:.1f} s)")

    # the product path: every contraction in FP64 by the O(N) semiseparable kernel
    b64s, ms64s, st64s, res64s = _scan_leg(scan, L_max, "fp64", "semiseparable", args.steps, args.warmup, graph,
                                           stream, tune=not args.no_tune)
    tuned = getattr(b64s, "tuning", None)
    log(f"fp64 semiseparable: {ms64s:.2f} ms/step, stages {st64s}, tuned {tuned.params if tuned else None}")
    mag = None if args.no_fp32 else b64s.magnitudes()
    sizes = {"alpha_beta_bytes": 0}
    for (ei, _, plan, *_rest) in b64s.pairs:
        # alpha (rows) and beta (cols) of both families, FP64
        sizes["alpha_beta_bytes"] += 2 * 8 * (plan.rows + plan.cols)
    ss_work = semisep_work_flop(b64s)
    dot_kind = getattr(b64s, "dot_kind", "tiled")
    kernel_timing = scan_kernel_ms(b64s, max(args.steps, 20), stream)
    log(f"scan kernel {kernel_timing[0]}: {kernel_timing[1]:.3f} ms")
    result = {
        "metric": METRIC, "value": n_int / (ms64s * 1e-3), "unit": "integrals/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms64s, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": (f"ionization-style batch: {len(scan)} ejected-energy points (k_e ~ U[0.2,4]) x "
                                f"{args.pairs} (l_i, l_f, l_ej <= 30) triples, Slater bound state (zeta 0.6) x 3 "
                                "continuum waves, per-point exp grid N~U{6000..15000}, L=0..30, direct+exchange"
                                if ionization else
                                "batched precision sweep: 64 energies x 400 channel pairs, per-energy exp grid "
                                "N~U{4000..15000} on [1e-5,1000] a0, L=0..30, direct+exchange, Full-Split "
                                "(FP64 alpha/beta prep and tails)")
                               + "; value = FP64 path with the O(N) semiseparable contraction; FP32-MFMA, "
                                 "split-BF16x3 and dense FP64 legs below",
                   "energies": len(scan), "pairs_per_energy": args.pairs, "L_max": L_max,
                   "integrals_per_step": n_int, "seed": args.seed, "graph": graph},
        "stage_ms": st64s,
        "autotune": ({"params": tuned.params, "ms": tuned.ms, "cached": tuned.cached, "candidates": tuned.timings}
                     if tuned else None),
        "step_tuning": ({"tail_fork": str(b64s.overlap), "direct_gemv": bool(b64s.direct_gemv),
                         "pipelined": bool(b64s.pipeline and b64s.pipeline_ok()),
                         "dot": b64s.dot_kind, **b64s.tuning_step}
                        if getattr(b64s, "tuning_step", None) else None),
        "host_note": "host planning, uploads and autotuning precede the timed region",
    }
    # stage rooflines: each stage's kernels timed on one stream, on the same batch
    stages_alone = stage_times(b64s)
    tail_ms = tail_entry_ms(b64s, max(args.steps, 10), stream) if b64s.n_tail else None
    W_tail, tail_nodes = tail_work_flop(b64s)
    p_bytes = prep_bytes(b64s)
    grouped = bool(b64s.prep_grouped and b64s.filon_groups is not None)
    wl = "ionization_48x400" if ionization else "batched_64x400"
    kname, k_ms = kernel_timing
    Wk = ss_work["scan_flop"]
    cands = {}
    t_p = stages_alone.get("prep")
    if t_p:
        tr, prov = _committed_traffic(["filon_fused_kernel", "filon_cells_kernel"] if grouped
                                      else ["prep_f64_kernel"], wl)
        cands["prep"] = {
            "bound": "hbm", "achieved": p_bytes / (t_p * 1e-3) / 1e9, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
            "frac": p_bytes / (t_p * 1e-3) / PEAK_HBM, "traffic": tr, "traffic_provenance": prov,
            "kernel": ("filon_fused_kernel + filon_cells_kernel + prep_f64_kernel (grouped Filon prep)" if grouped
                       else "prep_f64_kernel (per-pair Filon prep)"),
            "kernel_ms": t_p, "algorithmic_bytes_per_launch": p_bytes,
            "note": "prep stage alone on one stream; algorithmic bytes: chi_i, chi_f read once, alpha_dir, "
                    "alpha_ex, beta_ex written once per pair, shared grid data once per energy"}
    if tail_ms:
        # the tail kernels' multipole bucket (dwba_outer_tail_f64): smallest of these >= L_max + 1
        nl = next(v for v in (2, 4, 8, 16, 24, 31, 32, 48, 64) if L_max + 1 <= v)
        tk = [f"tail_filon_kernel<{nl}>", f"tail_levin_kernel<{nl}>", "tail_combine_kernel"]
        tr, prov = _committed_traffic(tk, wl)
        cands["tails"] = {
            "bound": "fp64_valu", "achieved": W_tail / (tail_ms * 1e-3) / 1e12, "peak": PEAK_FP64 / 1e12,
            "unit": "TFLOP/s", "frac": W_tail / (tail_ms * 1e-3) / PEAK_FP64, "traffic": tr,
            "traffic_provenance": prov,
            "kernel": " + ".join(tk) + " (dwba_outer_tail_f64)",
            "kernel_ms": tail_ms, "work_flop_per_launch": W_tail, "nodes_per_launch": tail_nodes,
            "note": "3 flop per swept node and multipole L = 1..L_max (one FMA into the sum, one multiply of "
                    "the x^-(L+1) chain); phases, weights and reciprocals are overhead"}
    if k_ms:
        tr, prov = _committed_traffic(kname, wl)
        cands["contraction"] = {
            "bound": "fp64_valu", "achieved": Wk / (k_ms * 1e-3) / 1e12, "peak": PEAK_FP64 / 1e12,
            "unit": "TFLOP/s", "frac": Wk / (k_ms * 1e-3) / PEAK_FP64, "traffic": tr, "traffic_provenance": prov,
            "kernel": kname, "kernel_ms": k_ms, "work_flop_per_launch": Wk,
            "algorithmic_bytes_per_launch": ss_work["scan_bytes"],
            "hbm_gbs_algorithmic": ss_work["scan_bytes"] / (k_ms * 1e-3) / 1e9,
            "note": "dominant contraction kernel timed alone; 9 flop per (grid point, L, family) of the "
                    "recurrences it runs; algorithmic bytes = the alpha and beta it reads, FP64, once"}
        t_c = stages_alone.get("plain")
        if t_c:
            W = ss_work["flop"]
            cands["contraction"]["contraction_stage"] = {
                "ms": t_c, "achieved_tflops": W / (t_c * 1e-3) / 1e12, "frac": W / (t_c * 1e-3) / PEAK_FP64,
                "work_flop": W,
                "kernels": kname + (" + semisep_gemv_scan_kernel + ddot_" + dot_kind + "_kernel"
                                    if ss_work["gemv"] else ""),
                "note": "all contraction kernels of a step (recurrences, shared-beta GEMV and per-pair dots at "
                        "2 flop per point and L)"}
    if cands:
        dom = max(cands, key=lambda k: cands[k]["kernel_ms"])
        result["roofline"] = dict(cands[dom], stage=dom,
                                  note_dominant="the stage with the longest kernel time alone; the others are in "
                                                "stage_rooflines")
        result["stage_rooflines"] = cands
    result["stage_ms_alone"] = stages_alone
    del b64s
    result["c_abi_scan"] = c_abi_scan(scan, L_max, args.steps, res64s, mag)
    log(f"C-ABI scan: {result['c_abi_scan']}")
    result["one_shot"] = one_shot_scan(scan, L_max)
    log(f"one-shot scan: {result['one_shot']}")
    if not ionization:
        result["dropin_latency"] = dropin_latency()
        log(f"drop-in: {result['dropin_latency']}")
        if not args.no_rows:
            result["continuum_waves"] = bench_continuum()
            log(f"continuum waves: {result['continuum_waves']}")
            result["partial_wave_sums"] = bench_partial_waves()
            log(f"partial-wave sums: {result['partial_wave_sums']}")

    if args.no_fp32:
        if cpu:
            result["cpu_baseline"] = cpu
        return result
    b32, ms32, st32, res32 = _scan_leg(scan, L_max, "fp32", "dense", args.steps, args.warmup, graph, stream)
    wd, we = b32.work_flop()
    del b32
    log(f"fp32 MFMA: {ms32:.1f} ms/step, stages {st32}")
    leg32 = {"value": n_int / (ms32 * 1e-3), "unit": "integrals/s", "ms_per_step": ms32, "stage_ms": st32,
             "work_flop": {"direct": wd, "exchange": we},
             "rel_err_vs_fp64": _err_stats(res32, res64s, mag)}
    if st32.get("xgemm"):
        ach = we / (st32["xgemm"] * 1e-3)
        leg32["roofline"] = {"bound": "mfma", "achieved": ach / 1e12, "peak": PEAK_FP32_MFMA / 1e12,
                             "unit": "TFLOP/s", "frac": ach / PEAK_FP32_MFMA,
                             "kernel": "xgemm_kernel<0> (v_mfma_f32_32x32x2_f32)", "kernel_ms": st32["xgemm"]}
    result["fp32_mfma"] = leg32
    del res32

    # block form of the same MFMA contraction: rank-one moments between blocks of
    # 8 points and each block's own 36 entries on the matrix cores
    for prec, key in (("fp32", "fp32_mfma_block"), ("bf16x3", "bf16x3_block")):
        if prec == "bf16x3" and args.no_bf16:
            continue
        bb, msb, stb, resb = _scan_leg(scan, L_max, prec, "block", args.steps, args.warmup, graph, stream)
        pp, fl = bb.block_work()
        del bb
        log(f"{prec} MFMA block form: {msb:.1f} ms/step, stages {stb}")
        leg = {"value": n_int / (msb * 1e-3), "unit": "integrals/s", "ms_per_step": msb, "stage_ms": stb,
               "rel_err_vs_fp64": _err_stats(resb, res64s, mag)}
        if stb.get("xgemm"):
            k_s = stb["xgemm"] * 1e-3
            peak = PEAK_FP32_MFMA if prec == "fp32" else 2.5e15
            work = fl if prec == "fp32" else 3 * fl
            leg["roofline"] = {
                "bound": "mfma", "achieved": work / k_s / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
                "frac": work / k_s / peak,
                "kernel": ("xblock_kernel<false> (v_mfma_f32_16x16x4_f32)" if prec == "fp32" else
                           "xblock_kernel<true> (v_mfma_f32_16x16x16_bf16, 3 products)") + " + xblock_join_kernel",
                "kernel_ms": stb["xgemm"], "work_flop": work, "pair_points": pp,
                "hbm_gbs_alpha_beta": 16.0 * pp / k_s / 1e9,
                "traffic": None, "traffic_provenance": None,
                "note": "MFMA work per (grid point, pair): four point moments (4 x 32 multipoles x 2 flop) and the "
                        "block's 36 entries / 8 points (32 x 36 x 2 / 8), 544 flop; split-BF16x3: 3 products of it; "
                        "hbm: FP64 alpha and beta read once"}
        if "roofline" in leg:
            kn = "xblock_kernel<false>" if prec == "fp32" else "xblock_kernel<true>"
            tr, prov = _committed_traffic([kn, "xblock_join_kernel"], "batched_64x400" if not ionization else
                                          "ionization_48x400")
            leg["roofline"]["traffic"], leg["roofline"]["traffic_provenance"] = tr, prov
        result[key] = leg
        del resb

    if not args.no_bf16:
        bbf, msbf, stbf, resbf = _scan_leg(scan, L_max, "bf16x3", "dense", args.steps, args.warmup, graph, stream)
        del bbf
        log(f"bf16x3: {msbf:.1f} ms/step, stages {stbf}")
        result["bf16x3"] = {"value": n_int / (msbf * 1e-3), "unit": "integrals/s", "ms_per_step": msbf,
                            "stage_ms": stbf,
                            "xgemm_tflops_effective": (we / (stbf["xgemm"] * 1e-3) / 1e12) if stbf.get("xgemm")
                            else None, "peak_effective_tflops": 2500.0 / 3,
                            "rel_err_vs_fp64": _err_stats(resbf, res64s, mag)}
        del resbf
    if not args.no_fp64:
        bd, msd, std, resd = _scan_leg(scan, L_max, "fp64", "dense", args.fp64_steps, 0, graph, stream)
        del bd
        log(f"fp64 dense plain: {msd:.1f} ms/step")
        result["fp64_dense"] = {"value": n_int / (msd * 1e-3), "unit": "integrals/s", "ms_per_step": msd,
                                "stage_ms": std, "kernel": "plain_f64_kernel (batched, FP64 VALU)",
                                "achieved_tflops": (wd + we) / (std.get("plain", msd) * 1e-3) / 1e12,
                                "max_abs_diff_vs_semiseparable_over_sum_abs_terms":
                                    float(np.max(np.abs(resd - res64s) / np.maximum(mag, 1e-300)))}
        result["speedups_vs_fp64_dense"] = {"fp64_semiseparable": msd / ms64s, "fp32_mfma": msd / ms32}
        for key in ("bf16x3", "fp32_mfma_block", "bf16x3_block"):
            if key in result:
                result["speedups_vs_fp64_dense"][key] = msd / result[key]["ms_per_step"]
    if cpu:
        result["cpu_baseline"] = cpu
    return result


def bench_continuum():
    """Continuum waves of one energy, both channels, as a scan computes them
    before its integrals (precompute_continuum_waves, driver.py:486-540):
    chi_i for l = 0..80 at 100 eV and chi_f for l = 0..84 at 89.8 eV on the
    15000-point exp grid of the batched scan ([1e-5, 1000] a0), H static
    potential, Johnson chain, hybrid phase.  Wall clock per solve call (host
    set-up, launches, copies back), with the analytic bypass on (the
    reference's default) and off (every wave propagated).  CPU baseline: the
    oracle's restatement of the same solver (one process) on a sample of
    propagated waves, scaled to the set."""
    import torch
    from dwba_amd.continuum import solve_continuum_waves
    from dwba_amd.grid import exp_grid
    from dwba_amd.synthetic import h_potentials
    gr = exp_grid(15000, 1e-5, 1e3)
    _, U = h_potentials(gr.r)
    sets = [(list(range(81)), 100.0), (list(range(85)), 89.8)]
    out = {}
    for byp in (True, False):
        for ls, E in sets:
            solve_continuum_waves(gr, U, ls, E, 0.0, "hybrid", "johnson", byp)       # warm
        torch.cuda.synchronize()
        reps = 3
        t0 = time.perf_counter()
        n_ok = 0
        for _ in range(reps):
            for ls, E in sets:
                ws = solve_continuum_waves(gr, U, ls, E, 0.0, "hybrid", "johnson", byp)
                n_ok += sum(w is not None for w in ws)
        dt = (time.perf_counter() - t0) / reps
        n = sum(len(ls) for ls, _ in sets)
        out["bypass_on" if byp else "bypass_off"] = {"waves": n, "s_per_energy": dt, "waves_per_s": n / dt,
                                                     "solved": n_ok // reps}
    from oracle import continuum as OC
    sample = [(0, 100.0), (40, 100.0), (80, 89.8)]
    t0 = time.perf_counter()
    for l, E in sample:
        OC.solve(gr.r, gr.w_trapz, U, l, E, 0.0, "hybrid", False)
    per = (time.perf_counter() - t0) / len(sample)
    n = out["bypass_off"]["waves"]
    out["cpu_baseline"] = {"value": 1.0 / per, "unit": "waves/s", "cores": 1, "kind": "port",
                           "sample": f"{len(sample)} propagated waves (l = 0, 40, 80), N = 15000, "
                                     f"{per:.2f} s each; the set of {n} would take {per * n:.0f} s"}
    out["speedup_vs_cpu_bypass_off"] = out["bypass_off"]["waves_per_s"] * per
    out["note"] = ("serial recursions, one lane per wave: latency-bound (no roofline applies); "
                   "waves/s of whole solve calls")
    return out


def bench_partial_waves():
    """Amplitudes, per-l_i sums, running totals and cross sections of 64
    energies on integrals already on the device (the partial-wave loop of
    driver.py:877-1152 after its integral calls): H 1s -> 2p (3 projection
    pairs), l_i <= 40, l_f <= l_i + 8 of odd l_i + l_f, L = 0..8, 200 angles;
    seeded synthetic integrals decaying with l.  Wall clock of
    excitation_sums_from_integrals.  CPU baseline: the oracle's restatement of
    the loop (reference amplitude routines) on one energy, scaled to 64."""
    import torch
    from dwba_amd import partial_waves as PW
    rng = np.random.default_rng(7)
    chan = PW.ExcitationChannel(0, 1, 8, 1.0)

    class _W:
        pass
    w = _W()
    ens = [PW.ExcitationEnergy(None, None, None, None, None, None, float(ki), float(kf),
                               {l: w for l in range(41)}, {l: w for l in range(49)}, 40)
           for ki, kf in rng.uniform(1.0, 4.0, (64, 2))]
    triples, has = PW.excitation_pairs(ens, chan)
    n = len(triples)
    I = np.zeros((n, 2, 9))
    for q, (_, li, lf) in enumerate(triples):
        I[q] = rng.standard_normal((2, 9)) * 0.7 ** (0.5 * (li + lf))
    dev = torch.as_tensor(I, device="cuda").contiguous()
    rows = np.arange(n)
    theta = np.linspace(0.0, np.pi, 200)
    PW.excitation_sums_from_integrals(dev, rows, triples, has, ens, chan, theta)      # warm
    torch.cuda.synchronize()
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        res = PW.excitation_sums_from_integrals(dev, rows, triples, has, ens, chan, theta)
    dt = (time.perf_counter() - t0) / reps
    from oracle import partial_waves as OP
    q0 = [q for q, t in enumerate(triples) if t[0] == 0]
    look = {(triples[q][1], triples[q][2]): q for q in q0}

    def integ(li, lf):
        q = look[(li, lf)]
        return {L: float(I[q, 0, L]) for L in range(9)}, {L: float(I[q, 1, L]) for L in range(9)}
    t0 = time.perf_counter()
    ref = OP.excitation_loop(theta, integ, lambda l: True, lambda l: l <= 48, ens[0].k_i, ens[0].k_f, 0, 1, 8, 40)
    tc = time.perf_counter() - t0
    agree = abs(res[0].sigma_au - ref["sigma_au"]) / max(abs(ref["sigma_au"]), 1e-300)
    return {"energies": 64, "pairs": n, "s_per_call": dt, "pairs_per_s": n / dt,
            "cpu_baseline": {"value": len(q0) / tc, "unit": "pairs/s", "cores": 1, "kind": "port",
                             "sample": f"one energy ({len(q0)} pairs, 3 projection pairs, 200 angles) "
                                       f"in {tc:.2f} s"},
            "speedup_vs_cpu": (n / dt) / (len(q0) / tc),
            "sigma_rel_diff_vs_oracle_energy0": agree,
            "note": "host pair tables, amplitude/sum/cumulate/DCS launches and the copy of the snapshots "
                    "back, per call"}


def one_shot_scan(scan, L_max):
    """What a scan that runs once pays, wall clock: host planning, uploads and
    device tables (ScanBatch construction), the launch choice (tune(): the
    per-device cache written by the timed leg, or library defaults), one eager
    launch with its synchronisation, and the copy of every I_L back into
    RadialDWBAIntegrals objects."""
    import torch
    from dwba_amd.batch import ScanBatch
    t0 = time.perf_counter()
    b = ScanBatch(scan, L_max, precision="fp64", oscillatory_method="full_split")
    t1 = time.perf_counter()
    b.tune()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    b.launch()
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    res = b.integrals()
    t4 = time.perf_counter()
    n = sum(len(r) for r in res)
    del b, res
    return {"construct_s": t1 - t0, "tune_s": t2 - t1, "launch_s": t3 - t2, "copy_back_s": t4 - t3,
            "total_s": t4 - t0, "pairs": n,
            "note": "fresh ScanBatch on the same scan: planning + uploads + tables, tune() (cached choice), "
                    "one eager launch, results to Python dicts"}


def c_abi_scan(scan, L_max, steps, ref=None, mag=None):
    """The same FP64 scan through the whole-scan C entry points (dwba_scan_plan
    / _launch / _capture / _replay, C++ planning and launch sequence), timed
    like the headline leg: graph replays on one stream between HIP events,
    inputs resident.  Also its host planning + upload time and the agreement
    with the ScanBatch result (|dI| / sum|a K b|)."""
    import torch
    from dwba_amd.cscan import CScan
    t0 = time.perf_counter()
    cs = CScan(scan, L_max, oscillatory_method="full_split")
    torch.cuda.synchronize()
    t_plan = time.perf_counter() - t0
    stream = torch.cuda.Stream()
    cs.launch(stream)
    stream.synchronize()
    eager_ms = _events_ms(lambda: cs.launch(stream), steps, stream)
    cs.capture(stream)
    cs.replay(stream)
    stream.synchronize()
    replay_ms = _events_ms(lambda: cs.replay(stream), steps, stream)
    res = cs.results(stream)
    info = cs.info()
    cs.close()
    n_int = 2 * res.shape[0] * res.shape[2]
    out = {"ms_per_step": replay_ms, "eager_ms_per_step": eager_ms, "value": n_int / (replay_ms * 1e-3),
           "unit": "integrals/s", "plan_s": t_plan, "info": info,
           "note": "include/dwba_hip.h dwba_scan_*: C++ host planning, one device allocation, fixed FP64 step "
                   "(tail fork, shared direct beta, split-lane contraction, grouped Filon prep)"}
    if ref is not None:
        d = np.abs(res - ref)
        out["max_abs_diff_vs_scanbatch"] = float(np.max(d))
        if mag is not None:
            out["max_diff_over_sum_abs_aKb"] = float(np.max(d / np.maximum(mag, 1e-300)))
    return out


def dropin_latency():
    """Per-call latency of the drop-in radial_ME_all_L_gpu over the 220 channel
    pairs of H 1s -> 2p at 50 eV (the reference's own inputs, tests/golden
    fixture: make_r_grid(1e-4, 200, 3000), waves l <= 20 from its solver,
    L = 0..4, full_split, CC 5, one GPUCache for the energy) -- the call
    sequence of driver.py:1016-1029 -- with the agreement against the
    reference's CPU values, and the same 220 pairs as one ScanBatch."""
    import torch
    from dwba_amd.batch import EnergyInput, ScanBatch
    from dwba_amd.dwba_matrix_elements import GPUCache, radial_ME_all_L_gpu
    from dwba_amd.synthetic import Bound, Wave
    path = os.path.join(ROOT, "tests", "golden", "c1_h1s2p_50eV.npz")
    if not os.path.exists(path):
        return None
    C1 = np.load(path)

    class _Grid:
        r, w_simpson, w_trapz = C1["r"], C1["w_simpson"], C1["w_trapz"]

    def wave(meta, chi):
        k, l, ps, im, eta, sig = meta
        return Wave(l=int(l), k_au=float(k), chi_of_r=chi, phase_shift=float(ps), idx_match=int(im),
                    eta=float(eta), sigma_l=float(sig))
    g = _Grid()
    bi = Bound(C1["u_i"], l=0, energy_au=float(C1["e_i"]))
    bf = Bound(C1["u_f"], l=1, energy_au=float(C1["e_f"]))
    wi = [wave(C1["meta_i"][l], C1["chi_i"][l]) for l in range(C1["chi_i"].shape[0])]
    wf = [wave(C1["meta_f"][l], C1["chi_f"][l]) for l in range(C1["chi_f"].shape[0])]
    pairs = [tuple(int(x) for x in p) for p in C1["pairs"]]
    Lm = int(C1["L_max"])
    gold = C1["I_full_split_cc5"]
    cache = GPUCache.from_grid(g)
    kw = dict(oscillatory_method="full_split", CC_nodes=5, gpu_cache=cache, U_f_array=C1["U_f"])
    radial_ME_all_L_gpu(g, C1["V_core"], C1["U_i"], bi, bf, wi[pairs[0][0]], wf[pairs[0][1]], Lm, **kw)   # warm
    times, worst = [], 0.0
    for q, (li, lf) in enumerate(pairs):
        t0 = time.perf_counter()
        res = radial_ME_all_L_gpu(g, C1["V_core"], C1["U_i"], bi, bf, wi[li], wf[lf], Lm, **kw)
        times.append(time.perf_counter() - t0)
        for fam, d in ((0, res.I_L_direct), (1, res.I_L_exchange)):
            for L in range(Lm + 1):
                ref = gold[q, fam, L]
                worst = max(worst, abs(d[L] - ref) / max(abs(ref), 1e-300))
    t = np.asarray(times) * 1e3
    en = EnergyInput(grid=g, V_core=C1["V_core"], U_i=C1["U_i"], bound_i=bi, bound_f=bf,
                     pairs=[(wi[li], wf[lf]) for (li, lf) in pairs], U_f=C1["U_f"])
    t0 = time.perf_counter()
    b = ScanBatch([en], Lm, precision="fp64", oscillatory_method="full_split", CC_nodes=5)
    b.launch()
    resb = b.results()
    t_once = time.perf_counter() - t0
    b.capture()
    b.run()
    ms_replay = _events_ms(b.run, 20, torch.cuda.current_stream())
    worst_b = float(np.max(np.abs(resb - gold) / np.maximum(np.abs(gold), 1e-300)))
    return {"calls": len(pairs), "mean_ms": float(t.mean()), "median_ms": float(np.median(t)),
            "p90_ms": float(np.percentile(t, 90)), "total_s": float(t.sum() / 1e3),
            "max_rel_vs_reference_cpu": worst, "scan_batch_one_shot_s": t_once, "scan_batch_replay_ms": ms_replay,
            "scan_batch_max_rel_vs_reference_cpu": worst_b,
            "workload": "H 1s->2p 50 eV, make_r_grid(1e-4,200,3000), 220 pairs l<=20, L=0..4, full_split CC5 "
                        "(reference inputs and CPU values from the tests/golden fixture)"}


_CPU_SCAN = None      # the scan the CPU-baseline workers read (set before the pool forks)


def _cpu_worker_init(threads):
    try:
        from threadpoolctl import threadpool_limits
        threadpool_limits(threads)
    except Exception:
        pass


def _cpu_faithful_call(task):
    """One faithful radial_ME_all_L restatement call (L = 0..Lm) on pair pi of
    energy ei of _CPU_SCAN; returns its wall time."""
    from oracle import radial
    ei, pi, Lm = task
    en = _CPU_SCAN[ei]
    pr = en.pairs[pi]
    bf = pr[2] if len(pr) > 2 else en.bound_f
    g = en.grid
    t0 = time.perf_counter()
    radial.radial_ME_all_L_faithful(g.r, g.w_simpson, g.w_trapz, en.V_core, en.U_i, en.bound_i, bf, pr[0], pr[1],
                                    Lm, oscillatory_method="full_split", U_f_array=en.U_f)
    return time.perf_counter() - t0


def cpu_baseline_batched(scan, L_max, workers=None):
    """The oracle's faithful restatement of radial_ME_all_L (materialises the
    N x N K_L of every call, Filon / filon_exchange node interpolations with
    N1 x n_cc gathers, per-L outer tails with math.fsum) in min(cpu_count, 8)
    worker processes, the reference's own CPU parallelism (driver.py:181-216,
    :1159-1247; one pair per task).  Sample: on the smallest-N and the
    largest-N energy, half the workers run one pair each with L = 0 and the
    other half the same pairs with L = 0..2, all at once; a call's cost is setup S + per-L P (from
    the two call lengths), modelled as a power of N between the two energies,
    and the whole batch is (sum over energies of pairs x (S + (L_max+1) P)) /
    workers.  Must run before the GPU is initialised (the pool forks)."""
    import multiprocessing as mp
    global _CPU_SCAN
    model, ncpu, blas = cpu_info()
    W = workers or min(ncpu or 1, 8)
    thr = max(1, (ncpu or 1) // W)
    _CPU_SCAN = scan
    Ns = [len(e.grid.r) for e in scan]
    picks = sorted({int(np.argmin(Ns)), int(np.argmax(Ns))})
    cost, walls = {}, {}
    ctx = mp.get_context("fork")
    with ctx.Pool(W, initializer=_cpu_worker_init, initargs=(thr,)) as pool:
        for e in picks:
            half = max(1, W // 2)         # the same pairs with L = 0 and with L = 0..2
            tasks = [(e, q % len(scan[e].pairs), 0) for q in range(half)] + \
                    [(e, q % len(scan[e].pairs), 2) for q in range(W - half)]
            t0 = time.perf_counter()
            ts = pool.map(_cpu_faithful_call, tasks, chunksize=1)
            walls[e] = time.perf_counter() - t0
            t_l0 = float(np.mean(ts[:half]))
            t_l2 = float(np.mean(ts[half:])) if W > half else t_l0
            P = max((t_l2 - t_l0) / 2.0, 1e-9)
            S = max(t_l0 - P, 0.0)
            cost[e] = (Ns[e], S, P)
    _CPU_SCAN = None
    (n_a, S_a, P_a), (n_b, S_b, P_b) = cost[picks[0]], cost[picks[-1]]
    expo = (lambda x, y: np.log(y / x) / np.log(n_b / n_a) if n_b != n_a else 2.0)
    pS, pP = expo(max(S_a, 1e-9), max(S_b, 1e-9)), expo(P_a, P_b)
    total = 0.0
    for e, x in enumerate(scan):
        S_e = S_a * (Ns[e] / n_a) ** pS
        P_e = P_a * (Ns[e] / n_a) ** pP
        total += len(x.pairs) * (S_e + (L_max + 1) * P_e)
    total /= W
    n_int = sum(len(x.pairs) for x in scan) * (L_max + 1) * 2
    return {"value": float(n_int / total), "unit": "integrals/s", "cores": W * thr, "kind": "port", "workers": W,
            "threads_per_worker": thr,
            "sample": (f"faithful radial_ME_all_L (full_split) restatement in {W} worker processes x {thr} BLAS "
                       f"threads: {W // 2} pairs with L=0 and the same with L=0..2, at once, on the N={n_a} "
                       f"and N={n_b} energies ({sum(walls.values()):.1f} s wall); per-call setup S and per-L cost "
                       f"P fitted as N^{pS:.2f} and N^{pP:.2f}; batch = sum pairs x (S + {L_max + 1} P) / {W}; "
                       f"host {model}, os.cpu_count={ncpu}"),
            "fit": {"N": [n_a, n_b], "setup_s": [S_a, S_b], "per_L_s": [P_a, P_b]}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=15000)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--workload", default="batched", choices=["batched", "ionization", "kernel_only", "backward"])
    ap.add_argument("--energies", type=int, default=64)
    ap.add_argument("--pairs", type=int, default=400)
    ap.add_argument("--fp64-steps", type=int, default=1)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-bf16", action="store_true")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32-MFMA (and BF16x3, FP64 dense) legs")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-rows", action="store_true", help="skip the continuum-wave and partial-wave-sum legs")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="library defaults instead of autotuned launch parameters")
    ap.add_argument("--copy-probe", action="store_true",
                    help="first copy 512 MiB device-to-device twice (calibration of HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the I_L path runs and is timed on a single MI355X; "
              "there is no multi-GPU mode to launch.", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        import torch
        x = torch.ones(64 << 20, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(2):
            y.copy_(x)
        torch.cuda.synchronize()
        print(f"[bench] copy probe: {x.numel() * 8} bytes read and written per copy", file=sys.stderr, flush=True)
        del x, y
    if args.workload == "ionization" and args.energies == 64:
        args.energies = 48          # the ionization workload has 48 ejected-energy points
    result = {"batched": bench_batched, "ionization": lambda a: bench_batched(a, ionization=True),
              "kernel_only": bench_kernel_only, "backward": bench_backward}[args.workload](args)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
