"""Batched scan engine: every (energy, channel pair) of a scan in one pass.

Each energy may have its own grid (LOCAL grids, driver.py / DW_main.py:1621-
1748).  For a batch the reference calls radial_ME_all_L(_gpu) once per pair
(driver.py:907-1029, ionization.py:526-608); here the whole scan is planned
once on the host and then runs as a fixed sequence of launches on one
stream, which can be captured into a HIP graph and replayed:

  1  dwba_fullsplit_prep_f64  alpha/beta of every pair (one workgroup each),
                              written straight into per-energy stacked matrices
  2  dwba_moments_f64         M_L per distinct bound-density pair
  3  dwba_outer_tail_f64      FP64 tails, one wavefront per (pair, L)
  4  direct family            shared beta -> dwba_dgemv_f64 + dwba_ddot_f64 (FP64);
                              distinct betas (ionization) -> stacked GEMM
  5  exchange family          precision "fp32": dwba_cast_f64_f32 + dwba_xgemm_f32
                              (MFMA); precision "fp64": both families through
                              dwba_il_semisep_f64 (O(N) per L, default) or the
                              dense dwba_il_plain_f64_batch
  6  dwba_batch_finalize      tails and L=0 corrections, in place

Pairs in the oscillatory-off branch (PLAIN mode) always take the FP64 plain
kernel for both families.
"""
from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from . import planner, xblock
from .planner import GridStats, PREP_PLAIN, PREP_FILON, PREP_STANDARD, PREP_ZERO

PREP_NO_ZERO_FILL = 0x10   # include/dwba_hip.h: the batch's per-pair vectors are zero-initialised once
PREP_FILON_GROUPED = 0x20  # include/dwba_hip.h: Filon alpha/beta from the group cell sums instead
PREP_NODES_SHARED = 0x40   # include/dwba_hip.h: Coulomb phase nodes filled by another job of the launch
PREP_FINITE_TAIL = 0x80    # include/dwba_hip.h: densities finite beyond N1 (the fused pass may stop there)
GROUP_SLICE = 4            # members per workgroup of the fused grouped Filon pass (its FUSE_MAX)
from .dwba_matrix_elements import (PairPlan, RadialDWBAIntegrals, TAIL_NONE, TAIL_OUTER, TAIL_LEGACY, plan_pair,
                                   _fit)

XROWS, XCOLS, XK = 256, 128, 32


class XgemmEnergy(ctypes.Structure):
    """dwba_xgemm_energy of include/dwba_hip.h."""
    _fields_ = [("logr", ctypes.c_void_p), ("B", ctypes.c_void_p), ("Bh", ctypes.c_void_p), ("Bl", ctypes.c_void_p),
                ("A", ctypes.c_void_p),
                ("col_extent", ctypes.c_void_p), ("row_blocks", ctypes.c_void_p), ("part_off", ctypes.c_int64),
                ("N", ctypes.c_int32), ("P", ctypes.c_int32), ("ld", ctypes.c_int32), ("n_cb", ctypes.c_int32),
                ("n_rb", ctypes.c_int32), ("pair_base", ctypes.c_int32), ("out_slot", ctypes.c_int32),
                ("pad_", ctypes.c_int32)]


class DgemvEnergy(ctypes.Structure):
    _fields_ = [("r", ctypes.c_void_p), ("beta", ctypes.c_void_p), ("t", ctypes.c_void_p),
                ("N", ctypes.c_int32), ("rows", ctypes.c_int32), ("ld", ctypes.c_int32), ("L_max", ctypes.c_int32)]


class LaneGrid(ctypes.Structure):
    """dwba_lane_grid of include/dwba_hip.h."""
    _fields_ = [("r", ctypes.c_void_p), ("tab", ctypes.c_void_p), ("n", ctypes.c_int32), ("pad_", ctypes.c_int32)]


class FilonGroup(ctypes.Structure):
    """dwba_filon_group of include/dwba_hip.h."""
    _fields_ = [("coef", ctypes.c_void_p), ("ld_coef", ctypes.c_int32), ("rep", ctypes.c_int32),
                ("first", ctypes.c_int32), ("count", ctypes.c_int32)]


class DdotPair(ctypes.Structure):
    _fields_ = [("alpha", ctypes.c_void_p), ("t", ctypes.c_void_p), ("n", ctypes.c_int32), ("ld", ctypes.c_int32),
                ("out_index", ctypes.c_int32), ("pad_", ctypes.c_int32)]


class FinalizePair(ctypes.Structure):
    _fields_ = [("scalars", ctypes.c_void_p), ("tails", ctypes.c_void_p), ("out_index", ctypes.c_int32),
                ("oscillatory", ctypes.c_int32), ("tail_kind", ctypes.c_int32), ("pad_", ctypes.c_int32)]


@dataclass
class EnergyInput:
    """One energy point of a scan: its grid, potentials and channel pairs.

    pairs: list of (cont_i, cont_f) or (cont_i, cont_f, bound_f) -- the third
    entry overrides bound_f per pair (ionization: the ejected wave)."""
    grid: object
    V_core: np.ndarray
    U_i: np.ndarray
    bound_i: object
    bound_f: object
    pairs: list
    U_f: Optional[np.ndarray] = None


def _last_nonfinite(a) -> int:
    bad = np.flatnonzero(~np.isfinite(a))
    return int(bad[-1]) if bad.size else -1


def _round_up(x, m):
    return (x + m - 1) // m * m


@dataclass
class _EnergyState:
    N: int
    ld: int
    gs: GridStats
    r_d: object
    logr_d: object
    w_d: object
    plans: list = field(default_factory=list)
    order: list = field(default_factory=list)        # batch pair ids of this energy in column order


class ScanBatch:
    """Plans and runs all I_L of a scan.  precision: "fp64" (every
    contraction in FP64), "fp32" (stacked contractions on FP32 MFMA) or
    "bf16x3" (split-BF16x3 MFMA with FP32 accumulation).  contraction: for
    FP64 "semiseparable" (O(N) per L, dwba_il_semisep_f64, default) or
    "dense" (O(N^2) per L, dwba_il_plain_f64_batch); for the MFMA modes
    "block" (dwba_xblock_contract: rank-one moments between blocks of 8
    points and each block's own entries on the matrix cores, L_max <= 31,
    the default there) or "dense" (dwba_xgemm_*: every K_L tile on the
    matrix cores, the default above L_max = 31).
    blocks_per_range: blocks of 8 points per workgroup of the block form."""

    def __init__(self, energies: Sequence[EnergyInput], L_max: int, precision: str = "fp64",
                 use_oscillatory_quadrature=True, oscillatory_method="advanced", CC_nodes=5,
                 phase_increment=1.5708, min_grid_fraction=0.10, k_threshold=0.5,
                 contraction: Optional[str] = None, blocks_per_range: int = 64):
        if precision not in ("fp32", "bf16x3", "fp64"):
            raise ValueError(f"precision must be 'fp32', 'bf16x3' or 'fp64', got {precision!r}")
        allowed = ("semiseparable", "dense") if precision == "fp64" else ("block", "dense")
        if contraction is None:             # the block form takes L_max <= 31
            contraction = allowed[0] if (precision == "fp64" or L_max <= 31) else "dense"
        if contraction not in allowed:
            raise ValueError(f"contraction for precision {precision!r} must be one of {allowed}, got {contraction!r}")
        self.contraction = contraction
        if L_max < 0 or L_max > 63:
            raise ValueError("L_max must be in [0, 63]")
        if contraction == "block" and L_max > 31:
            raise ValueError("the block-form MFMA contraction takes L_max <= 31; use contraction='dense'")
        self.blocks_per_range = int(blocks_per_range)
        import torch
        from . import _hip
        self.torch, self.hip = torch, _hip
        self.lib = _hip.load_library()
        self.dev = torch.device("cuda", _hip.ensure_init())
        self.L_max = int(L_max)
        self.precision = precision
        self.CC_nodes = int(CC_nodes)
        self._keep = []
        # defaults of the FP64 contraction options (the FP64 build and tune() set them)
        self.direct_gemv = False
        self.pipeline = False
        self._pipe = None
        self._side2 = None
        t0 = time.perf_counter()
        self._plan(energies, use_oscillatory_quadrature, oscillatory_method, phase_increment,
                   min_grid_fraction, k_threshold)
        self.plan_seconds = time.perf_counter() - t0
        t0 = time.perf_counter()
        self._build()
        torch.cuda.synchronize()
        self.build_seconds = time.perf_counter() - t0
        self.graph = None
        # "after_prep": the tail chain (moments, outer tails) runs on a second
        # stream, forked after prep and joined before finalize, so it overlaps
        # the contractions; "start": forked before prep; None: one stream.
        self.overlap = "after_prep"
        self._side = None
        self._side2 = None
        self.tuning_overlap = None

    # ------------------------------------------------------------ planning
    def _plan(self, energies, use_osc, method, phase_inc, mgf, k_thr):
        torch = self.torch
        T = lambda a: torch.as_tensor(a, device=self.dev)
        self.E = []
        self.pairs = []        # (energy index, pair index in energy, PairPlan, inputs)
        self._uploads = {}
        self._nf = {}
        self._bound_objs = [en.bound_i for en in energies]
        for ei, en in enumerate(energies):
            r = np.ascontiguousarray(en.grid.r, dtype=np.float64)
            N = len(r)
            wt = en.grid.w_trapz if hasattr(en.grid, "w_trapz") else np.gradient(r)
            gs = GridStats.of(r, wt)
            ld = _round_up(N, XK)
            st = _EnergyState(N=N, ld=ld, gs=gs, r_d=T(r), logr_d=T(np.log(np.maximum(r, 1e-300))),
                              w_d=T(np.ascontiguousarray(en.grid.w_simpson, dtype=np.float64)))
            V = _fit("V_core", en.V_core, N)
            Ui = _fit("U_i", en.U_i, N)
            Uf = _fit("U_f", en.U_f, N) if en.U_f is not None else None
            st.vdiff_d = T(V - Ui)
            st.ui = _fit("bound_i.u", en.bound_i.u_of_r, N)
            st.grid_nf = max(_last_nonfinite(V - Ui), _last_nonfinite(st.ui))
            ext = {}
            for pi, pr in enumerate(en.pairs):
                ci, cf = pr[0], pr[1]
                bf = pr[2] if len(pr) > 2 else en.bound_f
                uf = _fit("bound_f.u", bf.u_of_r, N)
                is_exc = not hasattr(bf, "k_au")
                plan = plan_pair(gs, Ui, Uf, st.ui, uf, ci, cf, self.L_max, use_osc, method, phase_inc, mgf,
                                 k_thr, is_exc, ext)
                self.pairs.append((ei, pi, plan, ci, cf, bf))
            self.E.append(st)

    def _nonfinite(self, obj, attr, N):
        """Last index < N where obj.attr is not finite (-1: none), once per object."""
        key = (id(obj), attr, N)
        hit = self._nf.get(key)
        if hit is not None and hit[0] is obj:
            return hit[1]
        v = _last_nonfinite(_fit(attr, getattr(obj, attr), N))
        self._nf[key] = (obj, v)
        return v

    def _dev_array(self, obj, attr, N):
        """Upload a wave / bound density once per object."""
        key = (id(obj), attr, N)
        hit = self._uploads.get(key)
        if hit is not None and hit[0] is obj:
            return hit[1]
        t = self.torch.as_tensor(_fit(attr, getattr(obj, attr), N), device=self.dev)
        self._uploads[key] = (obj, t)
        return t

    # ------------------------------------------------------------ device setup
    def _build(self):
        torch, hip, lib = self.torch, self.hip, self.lib
        dev, Lm = self.dev, self.L_max
        nl = Lm + 1
        f64 = torch.float64
        n_pairs = len(self.pairs)
        # global output order: PLAIN pairs first (FP64 plain batch), then per
        # energy in column order (pairs sorted by exchange support, descending)
        if self.precision == "fp64":
            plain_ids = []
            order = list(range(n_pairs))
            for ei, st in enumerate(self.E):
                st.order = [k for k in order if self.pairs[k][0] == ei]
        else:
            plain_ids = [k for k, p in enumerate(self.pairs) if p[2].mode == PREP_PLAIN]
            for ei, st in enumerate(self.E):
                ids = [k for k, p in enumerate(self.pairs) if p[0] == ei and p[2].mode != PREP_PLAIN]
                ids.sort(key=lambda k: -self._ex_support(self.pairs[k][2]))
                st.order = ids
            order = plain_ids + [k for st in self.E for k in st.order]
        self.out_index = np.empty(n_pairs, dtype=np.int64)
        self.out_index[order] = np.arange(n_pairs)
        self.n_pairs = n_pairs
        self.out = torch.zeros((n_pairs, 2, nl), dtype=f64, device=dev)

        # per-energy stacked alpha/beta (pairs in column order), zero padded
        cols = {}
        for ei, st in enumerate(self.E):
            st.members = st.order
            P = len(st.members)
            st.ad = torch.zeros((max(P, 1), st.ld), dtype=f64, device=dev)
            st.bd = torch.zeros((max(P, 1), st.ld), dtype=f64, device=dev)
            st.ae = torch.zeros((max(P, 1), st.ld), dtype=f64, device=dev)
            st.be = torch.zeros((max(P, 1), st.ld), dtype=f64, device=dev)
            for c, k in enumerate(st.members):
                cols[k] = (ei, c)
        # plain-mode pairs get private vectors
        self._plain_vec = {}
        for k in plain_ids:
            ei = self.pairs[k][0]
            if self.precision != "fp64":
                self._plain_vec[k] = torch.zeros((4, self.E[ei].ld), dtype=f64, device=dev)

        self.scalars = torch.zeros((n_pairs, 4), dtype=f64, device=dev)
        self.tails = torch.zeros((n_pairs, nl), dtype=f64, device=dev)

        # ---- prep jobs
        PJ = hip.PrepJob * n_pairs
        prep = PJ()
        coul_par = np.zeros((n_pairs, 4))
        any_coul = False
        node_bufs = []
        node_owner = {}
        self.mom_sets = {}   # (energy, id(u_f)) -> index
        for k, (ei, pi, plan, ci, cf, bf) in enumerate(self.pairs):
            st = self.E[ei]
            ui_d = self._dev_array(self._bound_i(ei), "u_of_r", st.N)
            uf_d = self._dev_array(bf, "u_of_r", st.N)
            ci_d = self._dev_array(ci, "chi_of_r", st.N)
            cf_d = self._dev_array(cf, "chi_of_r", st.N)
            if k in self._plain_vec:
                v = self._plain_vec[k]
                outs = (v[0], v[1], v[2], v[3])
            else:
                e2, c = cols[k]
                outs = (st.ad[c], st.bd[c], st.ae[c], st.be[c])
            lay = plan.layout
            nodes = None
            flags = PREP_NO_ZERO_FILL if plan.mode == PREP_FILON else 0
            if plan.mode == PREP_FILON and max(st.grid_nf, self._nonfinite(ci, "chi_of_r", st.N),
                                               self._nonfinite(cf, "chi_of_r", st.N),
                                               self._nonfinite(bf, "u_of_r", st.N)) < plan.N1:
                flags |= PREP_FINITE_TAIL
            if plan.mode == PREP_FILON and lay.coulomb:
                # pairs of one energy with the same stepping share one node
                # buffer: the first fills it, the others only count their nodes
                key = (ei, lay.p_start, lay.p_end, lay.p_step, lay.k_total, lay.eta_total)
                nodes = node_owner.get(key)
                if nodes is None:
                    nodes = torch.empty(10002, dtype=f64, device=dev)
                    node_owner[key] = nodes
                    node_bufs.append(nodes)
                else:
                    flags |= PREP_NODES_SHARED
                coul_par[k] = (lay.p_start, lay.p_end, lay.k_total, lay.eta_total)
                any_coul = True
            prep[k] = hip.PrepJob(hip.ptr(st.r_d), hip.ptr(st.logr_d), hip.ptr(st.w_d), hip.ptr(ui_d),
                                  hip.ptr(uf_d), hip.ptr(ci_d), hip.ptr(cf_d), hip.ptr(st.vdiff_d), hip.ptr(nodes),
                                  hip.ptr(outs[0]), hip.ptr(outs[1]), hip.ptr(outs[2]), hip.ptr(outs[3]),
                                  hip.ptr(self.scalars[k]), lay.p_start, lay.p_step, lay.p_end, st.N, plan.N1,
                                  lay.n_pnodes, plan.mode | flags, plan.log1, plan.log2)
        self.prep_d = hip.struct_array_to_device(prep, dev)
        self._prep_host = prep
        self.prep_d_gemv = None           # prep jobs for the shared-beta direct path (_build_fp64)
        self._gemv_nobeta = set()         # pairs whose beta_dir that path does not read
        self._prep_dev = {}               # (shared-beta path, grouped Filon) -> device prep jobs
        self._build_filon_groups()
        self.coul_d = torch.as_tensor(coul_par.ravel(), device=dev) if any_coul else None
        self.any_coul = any_coul
        self._keep += node_bufs
        cx, cw = planner.cc_reference(self.CC_nodes)
        self.cc = (np.ascontiguousarray(cx), np.ascontiguousarray(cw))

        # ---- moments and tails
        tail_ids = [k for k, p in enumerate(self.pairs) if p[2].tail_kind != TAIL_NONE]
        # longest first (each job writes its own row, so the order is free): the
        # wavefronts with thousands of Filon segments start before the short ones
        seg_max, dphi_max = planner.filon_limits()
        self.tail_nodes = {}
        for k in tail_ids:
            pl = self.pairs[k][2]
            self.tail_nodes[k] = planner.tail_sweep_nodes(pl.k_i, pl.l_i, pl.eta_i, pl.k_f, pl.l_f, pl.eta_f,
                                                          pl.r_match, pl.r_max, seg_max, dphi_max)
        tail_ids.sort(key=lambda k: -sum(self.tail_nodes[k]))
        self.n_tail = len(tail_ids)
        if tail_ids:
            sets, set_of = [], {}
            for k in tail_ids:
                ei, _, plan, ci, cf, bf = self.pairs[k]
                key = (ei, id(bf))
                if key not in set_of:
                    set_of[key] = len(sets)
                    sets.append((ei, bf))
            ns = len(sets)
            self.moments = torch.zeros((ns, nl), dtype=f64, device=dev)
            ptr_tab = np.zeros((4, ns), dtype=np.int64)
            n_tab = np.zeros(ns, dtype=np.int32)
            for s_i, (ei, bf) in enumerate(sets):
                st = self.E[ei]
                ptr_tab[:, s_i] = (hip.ptr(st.r_d), hip.ptr(st.w_d),
                                   hip.ptr(self._dev_array(self._bound_i(ei), "u_of_r", st.N)),
                                   hip.ptr(self._dev_array(bf, "u_of_r", st.N)))
                n_tab[s_i] = st.N
            self.mom_ptr = torch.as_tensor(ptr_tab, device=dev)
            self.mom_n = torch.as_tensor(n_tab, device=dev)
            self.n_mom_sets = ns
            TJ = hip.TailJob * len(tail_ids)
            tj = TJ()
            for q, k in enumerate(tail_ids):
                ei, _, plan, ci, cf, bf = self.pairs[k]
                s_i = set_of[(ei, id(bf))]
                tj[q] = hip.TailJob(plan.k_i, plan.delta_i, plan.eta_i, plan.sigma_i, plan.k_f, plan.delta_f,
                                    plan.eta_f, plan.sigma_f, plan.l_i, plan.l_f, plan.r_match, plan.r_max,
                                    hip.ptr(self.moments[s_i]), hip.ptr(self.tails[k]), Lm, plan.tail_kind)
            self.tail_d = hip.struct_array_to_device(tj, dev)
            self.tail_ws_bytes = lib.dwba_outer_tail_workspace_bytes(len(tail_ids), Lm)
            self.tail_ws = torch.empty(self.tail_ws_bytes, dtype=torch.uint8, device=dev)
            D, x = planner.levin_tables(8)
            self.levin = (np.ascontiguousarray(D.ravel()), np.ascontiguousarray(x))

        # ---- finalize table
        FJ = FinalizePair * n_pairs
        fj = FJ()
        for k, (ei, pi, plan, ci, cf, bf) in enumerate(self.pairs):
            fj[k] = FinalizePair(hip.ptr(self.scalars[k]), hip.ptr(self.tails[k]), int(self.out_index[k]),
                                 int(plan.oscillatory), plan.tail_kind, 0)
        self.fin_d = hip.struct_array_to_device(fj, dev)

        # ---- contraction tables
        if self.precision == "fp64":
            self._build_fp64(order)
        else:
            self._build_mfma(plain_ids)

    def _bound_i(self, ei):
        return self._bound_objs[ei]

    def _ex_support(self, plan: PairPlan) -> int:
        if plan.mode == PREP_FILON:
            return min(plan.N1 + 1, plan.N)
        if plan.mode == PREP_STANDARD:
            return plan.N
        return 0

    def _build_fp64(self, order):
        hip = self.hip
        jobs, nr, nc = [], [], []
        arr = (hip.PlainJob * self.n_pairs)()
        nra = (ctypes.c_int32 * self.n_pairs)()
        nca = (ctypes.c_int32 * self.n_pairs)()
        pos = {k: c for st in self.E for c, k in enumerate(st.members)}
        # semiseparable mode: an energy whose pairs share one bound_f has one
        # direct beta (u_f u_i w) for its Filon/standard pairs, so their direct
        # family is t_L = K_L beta once per energy (O(N)) and a dot per pair
        shared_direct = set()
        beta_owner = set()                # the pair of each energy whose beta_dir the GEMV reads
        dg, dd, dd_counts, dg_energy = [], [], [], []
        max_rows = 0
        self.t_bufs = []
        if self.contraction == "semiseparable":
            for ei, st in enumerate(self.E):
                ks = [k for k in st.members if self.pairs[k][2].mode in (PREP_FILON, PREP_STANDARD)]
                if not ks or len({id(self.pairs[k][5]) for k in st.members}) != 1:
                    continue
                rows_d = max(self.pairs[k][2].N1 for k in ks)
                dd_counts.append(len(ks))
                dg_energy.append(ei)
                t = self.torch.zeros((self.L_max + 1, st.ld), dtype=self.torch.float64, device=self.dev)
                self.t_bufs.append(t)
                dg.append(DgemvEnergy(hip.ptr(st.r_d), hip.ptr(st.bd[pos[ks[0]]]), hip.ptr(t), st.N, rows_d, st.ld,
                                      self.L_max))
                beta_owner.add(ks[0])
                max_rows = max(max_rows, rows_d)
                for k in ks:
                    shared_direct.add(k)
                    dd.append(DdotPair(hip.ptr(st.ad[pos[k]]), hip.ptr(t), self.pairs[k][2].N1, st.ld,
                                       int(self.out_index[k]), 0))
        self.ss_gemv = None
        if dg:
            # tiles of <= 64 consecutive pairs of one energy for the tiled dot
            items = []
            first = 0
            for cnt in dd_counts:
                for q in range(0, cnt, 64):
                    items.append((first + q, min(64, cnt - q)))
                first += cnt
            items_d = self.torch.as_tensor(np.asarray(items, dtype=np.int32).reshape(-1, 2), device=self.dev)
            self._keep.append(items_d)
            self.ss_gemv = (hip.struct_array_to_device((DgemvEnergy * len(dg))(*dg), self.dev), len(dg),
                            hip.struct_array_to_device((DdotPair * len(dd))(*dd), self.dev), len(dd),
                            items_d, len(items))
            # tiles of <= 16 pairs for the matrix-core dot
            items16 = []
            self._item16_energy = []
            first = 0
            for cnt, ei in zip(dd_counts, dg_energy):
                for q in range(0, cnt, 16):
                    items16.append((first + q, min(16, cnt - q)))
                    self._item16_energy.append(ei)
                first += cnt
            self.items16 = (self.torch.as_tensor(np.asarray(items16, dtype=np.int32).reshape(-1, 2),
                                                 device=self.dev), len(items16))
        arr_split = (hip.PlainJob * self.n_pairs)()
        nca_split = (ctypes.c_int32 * self.n_pairs)()
        for g, k in enumerate(order):
            ei, pi, plan, ci, cf, bf = self.pairs[k]
            st = self.E[ei]
            c = pos[k]
            arr[g] = hip.PlainJob(hip.ptr(st.r_d), hip.ptr(st.ad[c]), hip.ptr(st.bd[c]), hip.ptr(st.ae[c]),
                                  hip.ptr(st.be[c]), plan.rows, plan.cols)
            skip_dir = k in shared_direct      # direct family through the shared-beta GEMV instead
            # exchange alone: alpha_ex and beta_ex of Filon / plain pairs vanish at
            # and beyond N1, so the O(N) scan stops there
            cols_split = plan.rows if (skip_dir and plan.mode in (PREP_FILON, PREP_PLAIN)) else plan.cols
            arr_split[g] = hip.PlainJob(hip.ptr(st.r_d), 0 if skip_dir else hip.ptr(st.ad[c]),
                                        0 if skip_dir else hip.ptr(st.bd[c]), hip.ptr(st.ae[c]), hip.ptr(st.be[c]),
                                        plan.rows, cols_split)
            nra[g], nca[g] = plan.rows, plan.cols
            nca_split[g] = cols_split
        self.plain_d = hip.struct_array_to_device(arr, self.dev)
        self.plain_d_split = hip.struct_array_to_device(arr_split, self.dev) if self.ss_gemv is not None else None
        if self.ss_gemv is not None:
            # with the shared-beta path only one pair per energy writes beta_dir
            self._gemv_nobeta = set(shared_direct) - beta_owner
            self.prep_d_gemv = self._prep_jobs(True, False)
        self._job_energy = [self.pairs[k][0] for k in order]     # grid of each job, in job order
        self._dg_energy = dg_energy
        self.pipeline = False             # two-half pipelined FP64 step (tune())
        self._pipe = self._pipe_plan()
        self._jobs_host = (arr, arr_split)                        # host copies, job order
        self._lane_tabs = {}                                      # lb -> per-grid tables
        self._lane_sets = {}                                      # (lb, split) -> jobs sorted longest first
        self.plain_n_split = (nra, nca_split)
        self.direct_gemv = False          # chosen by tune() when the shared-beta path exists
        self.dot_kind = "stage"           # per-pair dots of that path on the FP64 matrix cores: "stage" (operands
                                          # staged through LDS) or "mfma" (direct loads); "tiled": FP64 VALU
        self.plain_n = (nra, nca)
        if self.contraction == "semiseparable":
            self.semisep_config = (0, 0)      # (seg, lb): library defaults until tuned
            self.semisep_lanes = False        # lane-serial kernel instead (chosen by tune())
            self.lane_tables = True           # that kernel reads per-grid 1/r, q, q^(lb b) tables
            self.lane_order = "length"        # job order of that kernel: "length" or "energy" (tune())
            self.lane_split = 0               # and splits each range between two wavefronts (tune()): 0 no,
                                              # 4 the LDS-staged split kernel (energy order, padded groups),
                                              # 2 or 3 the occupancy (waves per SIMD) the split kernel targets
            self.plain_ws_bytes = self._semisep_ws_bytes(0)
        else:
            self.plain_ws_bytes = self.lib.dwba_il_plain_f64_batch_workspace_bytes(self.n_pairs, nra, nca,
                                                                                   self.L_max)
        self.plain_ws = self.torch.empty(max(self.plain_ws_bytes, 1), dtype=self.torch.uint8, device=self.dev)
        self.plain_count = self.n_pairs

    def _build_mfma(self, plain_ids):
        torch, hip, lib = self.torch, self.hip, self.lib
        dev, Lm = self.dev, self.L_max
        nl = Lm + 1
        f64 = torch.float64
        # plain-mode pairs: FP64 plain batch writing out[0:len(plain_ids)]
        self.plain_count = len(plain_ids)
        if plain_ids:
            arr = (hip.PlainJob * len(plain_ids))()
            nra = (ctypes.c_int32 * len(plain_ids))()
            nca = (ctypes.c_int32 * len(plain_ids))()
            for g, k in enumerate(plain_ids):
                ei, pi, plan, ci, cf, bf = self.pairs[k]
                v = self._plain_vec[k]
                arr[g] = hip.PlainJob(hip.ptr(self.E[ei].r_d), hip.ptr(v[0]), hip.ptr(v[1]), hip.ptr(v[2]),
                                      hip.ptr(v[3]), plan.rows, plan.cols)
                nra[g], nca[g] = plan.rows, plan.cols
            self.plain_d = hip.struct_array_to_device(arr, dev)
            self.plain_n = (nra, nca)
            self.plain_ws_bytes = lib.dwba_il_plain_f64_batch_workspace_bytes(len(plain_ids), nra, nca, Lm)
            self.plain_ws = torch.empty(self.plain_ws_bytes, dtype=torch.uint8, device=dev)
        # direct family: shared beta per energy -> dgemv + ddot; else stacked GEMM
        dg, dd, xg_fams = [], [], []
        dd_counts = []                    # dots per shared-beta energy (consecutive in dd)
        owners, shared_pairs = set(), set()
        max_rows = 0
        self.t_bufs = []
        for ei, st in enumerate(self.E):
            if not st.order:
                continue
            plans = [self.pairs[k][2] for k in st.order]
            bfs = {id(self.pairs[k][5]) for k in st.order}
            shared = len(bfs) == 1
            rows_d = max(p.N1 if p.mode in (PREP_FILON, PREP_STANDARD) else 0 for p in plans)
            base = int(self.out_index[st.order[0]])
            if shared:
                if rows_d > 0:
                    src = next(c for c, p in enumerate(plans) if p.mode in (PREP_FILON, PREP_STANDARD))
                    t = torch.zeros((nl, st.ld), dtype=f64, device=dev)
                    self.t_bufs.append(t)
                    dg.append(DgemvEnergy(hip.ptr(st.r_d), hip.ptr(st.bd[src]), hip.ptr(t), st.N, rows_d, st.ld, Lm))
                    max_rows = max(max_rows, rows_d)
                    n0 = len(dd)
                    owners.add(st.order[src])
                    for c, k in enumerate(st.order):
                        p = self.pairs[k][2]
                        if p.mode in (PREP_FILON, PREP_STANDARD):
                            dd.append(DdotPair(hip.ptr(st.ad[c]), hip.ptr(t), p.N1, st.ld, base + c, 0))
                            shared_pairs.add(k)
                    dd_counts.append(len(dd) - n0)
            else:
                sup_d = [p.N if p.mode in (PREP_FILON, PREP_STANDARD) else 0 for p in plans]
                rows = [p.N1 if p.mode in (PREP_FILON, PREP_STANDARD) else 0 for p in plans]
                xg_fams.append((ei, st.ad, st.bd, rows, sup_d, base, 0))
            sup_e = [self._ex_support(p) for p in plans]
            rows_e = [p.N1 if p.mode in (PREP_FILON, PREP_STANDARD) else 0 for p in plans]
            xg_fams.append((ei, st.ae, st.be, rows_e, sup_e, base, 1))
        self.dgemv = None
        if dg:
            arr = (DgemvEnergy * len(dg))(*dg)
            self.dgemv = (hip.struct_array_to_device(arr, dev), len(dg), max_rows)
        self.ddot = None
        if dd:
            arr = (DdotPair * len(dd))(*dd)
            self.ddot = (hip.struct_array_to_device(arr, dev), len(dd))
        self.xg_count = 0
        self.xb_plan = None
        self.b32 = []
        if self.contraction == "block":
            # the FP64 parts of the step as the FP64 path runs them: grouped Filon
            # prep, beta_dir written once per shared-beta energy, t_L by the O(N)
            # GEMV and the per-pair dots on the FP64 matrix cores
            self.prep_grouped = self.filon_groups is not None
            if dg:
                self._gemv_nobeta = shared_pairs - owners
                self.prep_d_gemv = self._prep_jobs(True, False)
                self.direct_gemv = True
            if dd and Lm <= 31:
                items16, first = [], 0
                for cnt in dd_counts:
                    for q in range(0, cnt, 16):
                        items16.append((first + q, min(16, cnt - q)))
                    first += cnt
                self.items16 = (torch.as_tensor(np.asarray(items16, dtype=np.int32).reshape(-1, 2), device=dev),
                                len(items16))
            # per-grid factor tables (grid only: written once here) and the
            # groups of 64 pairs of every stacked family
            tabs, grids_d = xblock.build_tables(hip, lib, torch, dev, [(st.r_d, st.N) for st in self.E],
                                                self.precision)
            fams = []
            for (ei, Amat, Bmat, rows, sup, base, slot) in xg_fams:
                ext = [max(int(a), int(b)) for a, b in zip(rows, sup)]
                fams.append((ei, Amat, Bmat, self.E[ei].ld, ext, base, slot))
            self.xb_plan = xblock.XBlockPlan(hip, lib, torch, dev, fams, tabs, self.precision,
                                             blocks_per_range=self.blocks_per_range)
            self.xb_partial = self.xb_plan.partial(lib, torch, dev)
            self._keep += [tabs, grids_d]
            torch.cuda.current_stream().synchronize()
            return
        # stacked GEMM descriptors and work list
        descs, items, costs = [], [], []
        part_off = 0
        max_cb = 1
        for fi, (ei, Amat, Bmat, rows, sup, base, slot) in enumerate(xg_fams):
            st = self.E[ei]
            P = len(rows)
            n_cb = (P + XCOLS - 1) // XCOLS
            col_ext = np.zeros(n_cb, dtype=np.int32)
            row_blk = np.zeros(n_cb, dtype=np.int32)
            for cb in range(n_cb):
                sl = slice(cb * XCOLS, min(P, (cb + 1) * XCOLS))
                col_ext[cb] = max(sup[sl]) if sup[sl] else 0
                row_blk[cb] = (max(rows[sl]) + XROWS - 1) // XROWS if rows[sl] else 0
            n_rb = int(row_blk.max()) if n_cb else 0
            self._check_factor_range(st, n_rb)
            if self.precision == "fp32":
                Bf = torch.zeros((max(P, 1), st.ld), dtype=torch.float32, device=dev)
                Bh = Bl = None
                self.b32.append((Bmat, Bf))
            else:
                Bf = None
                Bh = torch.zeros((max(P, 1), st.ld), dtype=torch.int16, device=dev)
                Bl = torch.zeros((max(P, 1), st.ld), dtype=torch.int16, device=dev)
                self.b32.append((Bmat, (Bh, Bl)))
            ce_d = torch.as_tensor(col_ext, device=dev)
            rb_d = torch.as_tensor(row_blk, device=dev)
            self._keep += [ce_d, rb_d]
            descs.append(XgemmEnergy(hip.ptr(st.logr_d), hip.ptr(Bf), hip.ptr(Bh), hip.ptr(Bl), hip.ptr(Amat),
                                     hip.ptr(ce_d), hip.ptr(rb_d), part_off, st.N, P, st.ld, n_cb, n_rb, base, slot, 0))
            part_off += int(lib.dwba_xgemm_partial_doubles(n_rb, n_cb, Lm))
            max_cb = max(max_cb, n_cb)
            for cb in range(n_cb):
                ncol = min(XCOLS, P - cb * XCOLS)
                for rb in range(int(row_blk[cb])):
                    cost = min(XROWS, st.N - rb * XROWS) * int(col_ext[cb]) * ((ncol + 31) // 32)
                    for L in range(nl):
                        items.append((fi, L, cb, rb))
                        costs.append(cost)
        self.xg_count = len(descs)
        self.xg_max_cb = max_cb
        if descs:
            arr = (XgemmEnergy * len(descs))(*descs)
            self.xg_d = hip.struct_array_to_device(arr, dev)
            order = np.argsort(-np.asarray(costs, dtype=np.int64), kind="stable")
            it = np.asarray(items, dtype=np.int32).reshape(-1, 4)[order] if items else np.zeros((0, 4), np.int32)
            self.items_d = torch.as_tensor(np.ascontiguousarray(it), device=dev)
            self.n_items = len(items)
            self.partial = torch.zeros(max(part_off, 1), dtype=f64, device=dev)
            self.counter = torch.zeros(1, dtype=torch.int32, device=dev)

    def _check_factor_range(self, st, n_rb):
        lr = np.log(st.gs.r)
        for rb in range(n_rb):
            a = rb * XROWS
            b = min(st.N, a + XROWS) - 1
            if a < st.N and self.L_max * (lr[b] - lr[a]) > 80.0:
                raise ValueError("grid too coarse for the FP32 factorised kernel (row block spans "
                                 f"r in [{st.gs.r[a]:.3g}, {st.gs.r[b]:.3g}]); use precision='fp64'")

    def _semisep_ws_bytes(self, seg):
        """Workspace of the semiseparable contraction for either job set (all
        families, or exchange-only for the shared-beta pairs)."""
        nra, nca = self.plain_n
        nb = self.lib.dwba_il_semisep_f64_workspace_bytes(len(nra), nra, nca, self.L_max, seg)
        if getattr(self, "plain_d_split", None) is not None:
            nra_s, nca_s = self.plain_n_split
            nb = max(nb, self.lib.dwba_il_semisep_f64_workspace_bytes(len(nra_s), nra_s, nca_s, self.L_max, seg))
        return nb

    def _lane_table_set(self, lb):
        """Per-energy tables of the table-driven lane kernel for multipole
        block lb (built on first use): device grid descriptors and the
        job -> table pointer array."""
        hit = self._lane_tabs.get(lb)
        if hit is not None:
            return hit
        torch, hip, lib = self.torch, self.hip, self.lib
        stride = lib.dwba_semisep_lane_table_stride(self.L_max, lb)
        if stride <= 0:
            raise ValueError(f"lane tables: bad (L_max, lb) = ({self.L_max}, {lb})")
        tabs, grids = [], (LaneGrid * len(self.E))()
        for ei, st in enumerate(self.E):
            t = torch.empty(st.N * stride, dtype=torch.float64, device=self.dev)
            tabs.append(t)
            grids[ei] = LaneGrid(hip.ptr(st.r_d), hip.ptr(t), st.N, 0)
        ent = (hip.struct_array_to_device(grids, self.dev), len(self.E), max(st.N for st in self.E), tabs)
        # the tables depend on the grids only: filled once here, read by every launch
        hip.check(self.lib.dwba_semisep_lane_tables_f64(ent[1], hip.ptr(ent[0]), ent[2], self.L_max, lb,
                                                        hip.stream_handle()), "dwba_semisep_lane_tables_f64")
        torch.cuda.current_stream().synchronize()
        self._lane_tabs[lb] = ent
        return ent

    def _build_filon_groups(self):
        """Filon pairs of one energy with the same split index, phase layout
        and interpolation kinds share their cell sums: groups for
        dwba_filon_cells_f64 / dwba_filon_fused_f64 (prep_grouped)."""
        torch, hip = self.torch, self.hip
        self.prep_grouped = False
        self.filon_groups = None
        self._grouped_pairs = set()
        gl = planner.filon_groups((ei, plan) for (ei, _, plan, *_r) in self.pairs)
        n_slices = sum((len(g) + GROUP_SLICE - 1) // GROUP_SLICE for g in gl)
        if not gl or n_slices > 65535:
            return
        lds = [_round_up(self.pairs[g[0]][2].N1, 32) for g in gl]
        coef = torch.empty(7 * sum(lds), dtype=torch.float64, device=self.dev)
        arr = (FilonGroup * len(gl))()
        members, off, first = [], 0, 0
        for i, (g, ld) in enumerate(zip(gl, lds)):
            arr[i] = FilonGroup(hip.ptr(coef) + 8 * off, ld, g[0], first, len(g))
            members += g
            off += 7 * ld
            first += len(g)
        # the fused pass takes slices of <= GROUP_SLICE members of a group (same
        # cell sums), one workgroup each over its grid; each energy's slices go
        # to one XCD (its grid arrays, u_f and cell tables stay in that L2),
        # energies balanced over the XCDs, longest first on each
        per_e = {}
        for i, g in enumerate(gl):
            G = arr[i]
            ei = self.pairs[g[0]][0]
            for q in range(0, G.count, GROUP_SLICE):
                c = min(GROUP_SLICE, G.count - q)
                per_e.setdefault(ei, []).append((FilonGroup(G.coef, G.ld_coef, G.rep, G.first + q, c),
                                                 self.E[ei].N * (10 + 6 * c)))
        units = [(sum(c for _, c in v), [sl for sl, _ in v]) for v in per_e.values()]
        sl = planner.xcd_runs(units, pad=FilonGroup(arr[0].coef, arr[0].ld_coef, arr[0].rep, 0, 0))
        if len(sl) > 65535:
            return
        slices = (FilonGroup * len(sl))(*sl)
        self.filon_groups = (hip.struct_array_to_device(arr, self.dev), len(gl),
                             torch.as_tensor(np.asarray(members, dtype=np.int32), device=self.dev),
                             max(self.pairs[g[0]][2].N1 for g in gl), coef,
                             hip.struct_array_to_device(slices, self.dev), len(sl))
        self._grouped_pairs = set(members)

    def _prep_jobs(self, gemv, grouped):
        """Device prep jobs: with gemv, beta_dir only where the shared-beta
        path reads it; with grouped, the per-pair prep leaves the grouped Filon
        pairs to dwba_filon_fused_f64 (alpha / beta, beta_dir, L = 0 sums)."""
        key = (bool(gemv), bool(grouped))
        if key == (False, False):
            return self.prep_d
        hit = self._prep_dev.get(key)
        if hit is not None:
            return hit
        hip = self.hip
        pj = (hip.PrepJob * self.n_pairs)()
        for k in range(self.n_pairs):
            pj[k] = self._prep_host[k]
            if gemv and k in self._gemv_nobeta:
                pj[k].beta_dir = 0
            if grouped and k in self._grouped_pairs:
                pj[k].mode = pj[k].mode | PREP_FILON_GROUPED
        d = hip.struct_array_to_device(pj, self.dev)
        self._prep_dev[key] = d
        return d

    def _pipe_plan(self):
        """Split of the energies into two halves of about equal prep work
        (pairs x grid points) for the pipelined step: prep of the second half
        runs beside the contractions of the first.  None when there is
        nothing to split."""
        nE = len(self.E)
        if nE < 2 or self.contraction != "semiseparable":
            return None
        cost = np.zeros(nE)
        for (ei, _, plan, *_r) in self.pairs:
            cost[ei] += self.E[ei].N
        cum = np.cumsum(cost)
        eA = int(np.searchsorted(cum, 0.5 * cum[-1])) + 1
        eA = min(max(eA, 1), nE - 1)
        kA = next((k for k, p in enumerate(self.pairs) if p[0] >= eA), len(self.pairs))
        dgA = sum(1 for e in getattr(self, "_dg_energy", []) if e < eA)
        itA = sum(1 for e in getattr(self, "_item16_energy", []) if e < eA)
        return {"eA": eA, "kA": kA, "dgA": dgA, "itA": itA}

    def _lane_job_set(self, lb, split, half=None, group=0):
        """Jobs of the table-driven lane kernel sorted by length, longest
        first (wavefronts of equal-length pairs, long ones dispatched first),
        with their table pointers and output rows; half 0 / 1: only the jobs
        of that half of the energies (pipelined step).  lane_order "energy":
        pairs of one energy kept together (energies with the longest pairs
        first, longest first within each), so a wavefront's grid-table reads
        hit one table.  group > 0 (the LDS-staged kernel): energy order with
        each energy's jobs padded to whole groups of that many by empty jobs
        whose output index is -1, so a group never spans two grid tables."""
        key = (lb, bool(split), half, "energy" if group else self.lane_order, group,
               getattr(self, "stage_layout", "xcd") if group else None)
        hit = self._lane_sets.get(key)
        if hit is not None:
            return hit
        torch, hip = self.torch, self.hip
        _, _, _, tabs = self._lane_table_set(lb)
        src = self._jobs_host[1 if split else 0]
        jobs = range(self.plain_count)
        if half is not None:
            eA = self._pipe["eA"]
            jobs = [g for g in jobs if (self._job_energy[g] >= eA) == (half == 1)]
        if self.lane_order == "energy" or group:
            emax = {}
            for g in jobs:
                e = self._job_energy[g]
                emax[e] = max(emax.get(e, 0), int(src[g].n_cols))
            perm = sorted(jobs, key=lambda g: (-emax[self._job_energy[g]], self._job_energy[g], -int(src[g].n_cols)))
        else:
            perm = sorted(jobs, key=lambda g: -int(src[g].n_cols))
        energy = [self._job_energy[g] for g in perm]
        if group and perm:
            padded, pe = [], []
            for p, g in enumerate(perm):
                padded.append(g)
                pe.append(energy[p])
                if p + 1 == len(perm) or energy[p + 1] != energy[p]:
                    while len(padded) % group:
                        padded.append(-1)
                        pe.append(energy[p])
            glen = lambda s: int(src[padded[s]].n_cols)   # a group's first job is its longest
            starts = list(range(0, len(padded), group))
            if getattr(self, "stage_layout", "xcd") == "xcd":
                # Workgroup w (two groups) runs on XCD w mod 8: each energy's groups go to one XCD, so its grid
                # table is read from that XCD's L2.  Energies are dealt to eight queues longest first (LPT on
                # points), each queue's groups longest first, two per workgroup, and the queues interleaved
                # (empty groups fill the short queues); the exchange family starts at a multiple of eight
                # workgroups since n_jobs becomes a multiple of 16 groups.
                by_e = {}
                for st in starts:
                    by_e.setdefault(pe[st], []).append(st)
                cost = {e: sum(glen(st) for st in v) for e, v in by_e.items()}
                qs, load = [[] for _ in range(8)], [0] * 8
                for e in sorted(by_e, key=lambda e: -cost[e]):
                    x = int(np.argmin(load))
                    load[x] += cost[e]
                    qs[x] += by_e[e]
                for q in qs:
                    q.sort(key=lambda st: -glen(st))
                    if len(q) % 2:
                        q.append(None)
                depth = max(len(q) for q in qs) // 2
                order = []
                for i in range(depth):
                    for q in qs:
                        order += q[2 * i:2 * i + 2] if 2 * i < len(q) else [None, None]
            else:
                order = sorted(starts, key=lambda st: -glen(st))      # longest first across energies
            perm = [padded[st + q] if st is not None else -1 for st in order for q in range(group)]
            e0 = pe[0]
            energy = [pe[st + q] if st is not None else e0 for st in order for q in range(group)]
        n = len(perm)
        arr = (hip.PlainJob * max(n, 1))()
        for p, g in enumerate(perm):
            if g >= 0:
                arr[p] = src[g]
            else:
                j = src[next(q for q in perm if q >= 0)]
                arr[p] = hip.PlainJob(j.r, j.alpha_dir, j.beta_dir, j.alpha_ex, j.beta_ex, 0, 0)
        jobs_d = hip.struct_array_to_device(arr, self.dev)
        ptrs = torch.as_tensor(np.asarray([hip.ptr(tabs[e]) for e in energy], dtype=np.int64), device=self.dev)
        oidx = torch.as_tensor(np.asarray(perm, dtype=np.int32), device=self.dev)
        ent = (jobs_d, ptrs, oidx, n)
        self._lane_sets[key] = ent
        return ent

    def _contract_semisep(self, seg, lb, direct_gemv, out, ws, ws_bytes, sh, lanes=False, split=None):
        """All FP64 contractions of the batch through the semiseparable kernel;
        with direct_gemv the Filon/standard pairs of shared-beta energies take
        their direct family from t_L = K_L beta (once per energy) and a dot."""
        hip, lib, Lm = self.hip, self.lib, self.L_max
        nra, nca = self.plain_n_split if direct_gemv else self.plain_n
        jobs = self.plain_d_split if direct_gemv else self.plain_d
        if lanes and self.lane_tables:
            lbv = lb or 8
            split = self.lane_split if split is None else split     # 0: whole ranges, 2 or 3: waves per SIMD,
            if split == 4:                                          # 4: LDS-staged
                sj, ptrs, oidx, n = self._lane_job_set(lbv, direct_gemv, group=self._stage_group(lbv))
                hip.check(lib.dwba_il_semisep_lanes_stage_f64(n, hip.ptr(sj), hip.ptr(ptrs), hip.ptr(oidx), Lm, lbv,
                                                              hip.ptr(out), sh), "dwba_il_semisep_lanes_stage_f64")
                return self._contract_direct(direct_gemv, out, sh)
            sj, ptrs, oidx, _ = self._lane_job_set(lbv, direct_gemv)
            if split:
                hip.check(lib.dwba_il_semisep_lanes_split_f64(self.plain_count, hip.ptr(sj), hip.ptr(ptrs),
                                                              hip.ptr(oidx), Lm, lbv, 3 if int(split) == 3 else 2,
                                                              hip.ptr(out), sh), "dwba_il_semisep_lanes_split_f64")
            else:
                hip.check(lib.dwba_il_semisep_lanes_tab_f64(self.plain_count, hip.ptr(sj), hip.ptr(ptrs),
                                                            hip.ptr(oidx), Lm, lbv, hip.ptr(out), sh),
                          "dwba_il_semisep_lanes_tab_f64")
        elif lanes:
            hip.check(lib.dwba_il_semisep_lanes_f64(self.plain_count, hip.ptr(jobs), Lm, lb, hip.ptr(out), sh),
                      "dwba_il_semisep_lanes_f64")
        else:
            hip.check(lib.dwba_il_semisep_f64(self.plain_count, hip.ptr(jobs), nra, nca, Lm, seg, lb, hip.ptr(out),
                                              hip.ptr(ws), ws_bytes, sh), "dwba_il_semisep_f64")
        self._contract_direct(direct_gemv, out, sh)

    def _stage_group(self, lb):
        """Jobs per group of the LDS-staged lane kernel (its wavefront pair's
        pairs) at lb multipoles per lane; 0 where it does not apply."""
        return max(0, int(self.lib.dwba_semisep_lane_stage_group(self.L_max, lb)))

    def _contract_direct(self, direct_gemv, out, sh):
        """The shared-beta direct family: t_L = K_L beta per energy, then the dots."""
        hip, lib, Lm = self.hip, self.lib, self.L_max
        if direct_gemv:
            gd, ng, dd, nd, items_d, ni = self.ss_gemv
            hip.check(lib.dwba_semisep_gemv_f64(ng, hip.ptr(gd), Lm, sh), "dwba_semisep_gemv_f64")
            if Lm <= 31 and self.dot_kind in ("mfma", "stage"):
                it16, n16 = self.items16
                fn = lib.dwba_ddot_stage_f64 if self.dot_kind == "stage" else lib.dwba_ddot_mfma_f64
                hip.check(fn(n16, hip.ptr(it16), hip.ptr(dd), Lm, hip.ptr(out), sh), "dwba_ddot_" + self.dot_kind + "_f64")
            elif Lm <= 31:
                hip.check(lib.dwba_ddot_tiled_f64(ni, hip.ptr(items_d), hip.ptr(dd), Lm, hip.ptr(out), sh),
                          "dwba_ddot_tiled_f64")
            else:
                hip.check(lib.dwba_ddot_f64(nd, hip.ptr(dd), Lm, hip.ptr(out), sh), "dwba_ddot_f64")

    def pipeline_ok(self):
        """The pipelined step needs the table-driven lane kernel and, with the
        shared-beta path, the matrix-core dots (per-half job, GEMV and item
        ranges exist for those)."""
        return (self._pipe is not None and self.precision == "fp64" and self.contraction == "semiseparable"
                and not self.prep_grouped and self.semisep_lanes and self.lane_tables and self.L_max <= 31
                and (not self.direct_gemv or self.dot_kind in ("mfma", "stage")))

    def _contract_half(self, half, out, sh):
        """The contraction of one half of the energies (pipelined step): the
        shared-beta GEMV and dots of its energies, then the lane kernel on
        its jobs.  The grid tables must already be written."""
        hip, lib, Lm = self.hip, self.lib, self.L_max
        seg, lb = self.semisep_config
        lbv = lb or 8
        P = self._pipe
        if self.direct_gemv:
            gd, ng, dd, nd, _, _ = self.ss_gemv
            it16, n16 = self.items16
            g0, gn = (0, P["dgA"]) if half == 0 else (P["dgA"], ng - P["dgA"])
            i0, i_n = (0, P["itA"]) if half == 0 else (P["itA"], n16 - P["itA"])
            if gn > 0:
                hip.check(lib.dwba_semisep_gemv_f64(gn, hip.ptr(gd) + g0 * ctypes.sizeof(DgemvEnergy), Lm, sh),
                          "dwba_semisep_gemv_f64")
            if i_n > 0:
                fn = lib.dwba_ddot_stage_f64 if self.dot_kind == "stage" else lib.dwba_ddot_mfma_f64
                hip.check(fn(i_n, hip.ptr(it16) + i0 * 8, hip.ptr(dd), Lm, hip.ptr(out), sh),
                          "dwba_ddot_" + self.dot_kind + "_f64")
        if int(self.lane_split) == 4:
            sj, ptrs, oidx, n = self._lane_job_set(lbv, self.direct_gemv, half, group=self._stage_group(lbv))
            if n:
                hip.check(lib.dwba_il_semisep_lanes_stage_f64(n, hip.ptr(sj), hip.ptr(ptrs), hip.ptr(oidx), Lm, lbv,
                                                              hip.ptr(out), sh), "dwba_il_semisep_lanes_stage_f64")
            return
        sj, ptrs, oidx, n = self._lane_job_set(lbv, self.direct_gemv, half)
        if n == 0:
            return
        if self.lane_split:
            hip.check(lib.dwba_il_semisep_lanes_split_f64(n, hip.ptr(sj), hip.ptr(ptrs), hip.ptr(oidx), Lm, lbv,
                                                          3 if int(self.lane_split) == 3 else 2, hip.ptr(out), sh),
                      "dwba_il_semisep_lanes_split_f64")
        else:
            hip.check(lib.dwba_il_semisep_lanes_tab_f64(n, hip.ptr(sj), hip.ptr(ptrs), hip.ptr(oidx), Lm, lbv,
                                                        hip.ptr(out), sh), "dwba_il_semisep_lanes_tab_f64")

    # ------------------------------------------------------------ launch
    def launch(self, stream=None, events=None):
        """Enqueue the whole scan on `stream` (torch's current by default).
        `events`: optional dict; HIP events are recorded around each stage
        under keys (stage, 'start'|'end') for timing on the launch stream."""
        hip, lib = self.hip, self.lib
        sh = hip.stream_handle(stream)
        torch = self.torch
        strm = stream if stream is not None else torch.cuda.current_stream()

        def mark(name, which):
            if events is not None:
                ev = torch.cuda.Event(enable_timing=True)
                ev.record(strm)
                events[(name, which)] = ev
        Lm = self.L_max

        def tails(h):
            if not self.n_tail:
                return
            hip.check(lib.dwba_moments_f64(self.n_mom_sets, hip.ptr(self.mom_ptr[0]), hip.ptr(self.mom_ptr[1]),
                                           hip.ptr(self.mom_ptr[2]), hip.ptr(self.mom_ptr[3]), hip.ptr(self.mom_n),
                                           Lm, hip.ptr(self.moments), h), "dwba_moments_f64")
            seg, dmax = planner.filon_limits()
            Dp, D_keep = hip.f64_array(self.levin[0])     # the arrays stay bound until the call returns
            Xp, X_keep = hip.f64_array(self.levin[1])
            hip.check(lib.dwba_outer_tail_f64(self.n_tail, hip.ptr(self.tail_d), Lm, seg, dmax, Dp, Xp,
                                              hip.ptr(self.tail_ws), self.tail_ws_bytes, h), "dwba_outer_tail_f64")

        # The tail chain reads only bound densities, phases and its own
        # moments -- nothing prep or the contractions write -- so without
        # stage events it runs on a side stream (default priority) forked
        # from the launch stream and joined before finalize.
        fork = events is None and bool(self.n_tail) and self.overlap in ("start", "after_prep")
        ev_join = None

        def fork_tails():
            nonlocal ev_join
            if self._side is None:
                self._side = torch.cuda.Stream(device=self.dev)
            ev_fork = torch.cuda.Event()
            ev_fork.record(strm)
            self._side.wait_event(ev_fork)
            tails(hip.stream_handle(self._side))
            ev_join = torch.cuda.Event()
            ev_join.record(self._side)

        mark("total", "start")
        if fork and self.overlap == "start":
            fork_tails()
        mark("prep", "start")
        cxp, cx_keep = hip.f64_array(self.cc[0])         # bound for the whole launch (prep() reads them)
        cwp, cw_keep = hip.f64_array(self.cc[1])
        gemv = getattr(self, "direct_gemv", False) and self.prep_d_gemv is not None
        grouped = self.prep_grouped and self.filon_groups is not None
        prep_jobs = self._prep_jobs(gemv, grouped)

        def prep(k0, n, h):
            if n <= 0:
                return
            coul = hip.ptr(self.coul_d) + 4 * 8 * k0 if self.any_coul else 0
            hip.check(lib.dwba_fullsplit_prep_f64(n, hip.ptr(prep_jobs) + k0 * ctypes.sizeof(hip.PrepJob), coul,
                                                  int(self.any_coul), cxp, cwp, self.CC_nodes, h),
                      "dwba_fullsplit_prep_f64")
            if grouped:                    # whole batch only (the pipelined step does not group)
                gd, ng, mem, max_n1, _, sd, ns = self.filon_groups
                hip.check(lib.dwba_filon_cells_f64(ng, hip.ptr(gd), hip.ptr(prep_jobs), cxp, cwp, self.CC_nodes, h),
                          "dwba_filon_cells_f64")
                hip.check(lib.dwba_filon_fused_f64(ns, hip.ptr(sd), hip.ptr(prep_jobs), hip.ptr(mem), h),
                          "dwba_filon_fused_f64")

        # Pipelined step (no stage events): prep of the second half of the
        # energies runs on the launch stream while the first half is
        # contracted on a second stream, which then takes the second half.
        pipe = events is None and self.pipeline and self.plain_count and self.pipeline_ok()
        if pipe:
            kA = self._pipe["kA"]
            if self._side2 is None:
                self._side2 = torch.cuda.Stream(device=self.dev)
            s2, sh2 = self._side2, hip.stream_handle(self._side2)
            prep(0, kA, sh)
            ev_a = torch.cuda.Event()
            ev_a.record(strm)
            if fork and self.overlap == "after_prep":
                fork_tails()
            prep(kA, self.n_pairs - kA, sh)
            ev_b = torch.cuda.Event()
            ev_b.record(strm)
            s2.wait_event(ev_a)
            self._contract_half(0, self.out, sh2)
            s2.wait_event(ev_b)
            self._contract_half(1, self.out, sh2)
            ev_c = torch.cuda.Event()
            ev_c.record(s2)
            strm.wait_event(ev_c)
        else:
            prep(0, self.n_pairs, sh)
        mark("prep", "end")
        if fork and self.overlap == "after_prep" and not pipe:
            fork_tails()
        mark("tail", "start")
        if not fork:
            tails(sh)
        mark("tail", "end")
        mark("plain", "start")
        if self.plain_count and not pipe:
            nra, nca = self.plain_n
            if self.contraction == "semiseparable":
                seg, lb = self.semisep_config
                self._contract_semisep(seg, lb, self.direct_gemv, self.out, self.plain_ws, self.plain_ws_bytes, sh,
                                       self.semisep_lanes)
            else:
                hip.check(lib.dwba_il_plain_f64_batch(self.plain_count, hip.ptr(self.plain_d), nra, nca, Lm,
                                                      hip.ptr(self.out), hip.ptr(self.plain_ws), self.plain_ws_bytes,
                                                      sh), "dwba_il_plain_f64_batch")
        mark("plain", "end")
        if self.precision in ("fp32", "bf16x3"):
            mark("direct", "start")
            if self.dgemv is not None:
                d, n, mr = self.dgemv
                if self.contraction == "block":     # t_L = K_L beta in O(N) per L, FP64
                    hip.check(lib.dwba_semisep_gemv_f64(n, hip.ptr(d), Lm, sh), "dwba_semisep_gemv_f64")
                else:
                    hip.check(lib.dwba_dgemv_f64(n, hip.ptr(d), mr, Lm, sh), "dwba_dgemv_f64")
            if self.ddot is not None:
                d, n = self.ddot
                if self.contraction == "block" and getattr(self, "items16", None) is not None:
                    it16, n16 = self.items16
                    hip.check(lib.dwba_ddot_stage_f64(n16, hip.ptr(it16), hip.ptr(d), Lm, hip.ptr(self.out), sh),
                              "dwba_ddot_stage_f64")
                else:
                    hip.check(lib.dwba_ddot_f64(n, hip.ptr(d), Lm, hip.ptr(self.out), sh), "dwba_ddot_f64")
            mark("direct", "end")
            if self.xb_plan is not None:
                mark("xgemm", "start")
                self.xb_plan.launch(hip, lib, self.precision, Lm, self.xb_partial, self.out, strm)
                mark("xgemm", "end")
            if self.xg_count:
                for Bm, Bx in self.b32:
                    if self.precision == "fp32":
                        hip.check(lib.dwba_cast_f64_f32(hip.ptr(Bm), hip.ptr(Bx), Bm.numel(), sh),
                                  "dwba_cast_f64_f32")
                    else:
                        hip.check(lib.dwba_split_bf16(hip.ptr(Bm), hip.ptr(Bx[0]), hip.ptr(Bx[1]), Bm.numel(), sh),
                                  "dwba_split_bf16")
                mark("xgemm", "start")
                fn = lib.dwba_xgemm_f32 if self.precision == "fp32" else lib.dwba_xgemm_bf16x3
                hip.check(fn(self.xg_count, hip.ptr(self.xg_d), self.n_items, hip.ptr(self.items_d),
                             Lm, hip.ptr(self.partial), hip.ptr(self.counter), self.xg_max_cb,
                             hip.ptr(self.out), 0, sh), "dwba_xgemm")
                mark("xgemm", "end")
        if ev_join is not None:
            strm.wait_event(ev_join)
        hip.check(lib.dwba_batch_finalize(self.n_pairs, hip.ptr(self.fin_d), Lm, hip.ptr(self.out), sh),
                  "dwba_batch_finalize")
        mark("total", "end")

    def stage_ms(self, events):
        """Stage durations (ms) from the events recorded by launch(events=...)."""
        out = {}
        for (name, which), ev in events.items():
            if which == "start" and (name, "end") in events:
                out[name] = ev.elapsed_time(events[(name, "end")])
        return out

    def magnitudes(self):
        """sum_ij |alpha_i| K_L(i,j) |beta_j| per (pair, family, L) in FP64: the
        cancellation-free size of each contraction, used to state errors of
        reduced-precision modes relative to the terms that are summed.  Needs
        the stacked alpha/beta of a run (call after run()).  Returns an array
        in the order the pairs were given."""
        torch, hip, lib = self.torch, self.hip, self.lib
        n = self.n_pairs
        if (getattr(self, "direct_gemv", False) and self.prep_d_gemv is not None) or self.prep_grouped:
            # the shared-beta path leaves beta_dir of most pairs unwritten: write all
            # (the per-pair prep also recomputes grouped pairs' alpha / beta_ex)
            cxp, cx_keep = hip.f64_array(self.cc[0])
            cwp, cw_keep = hip.f64_array(self.cc[1])
            hip.check(lib.dwba_fullsplit_prep_f64(n, hip.ptr(self.prep_d), hip.ptr(self.coul_d), int(self.any_coul),
                                                  cxp, cwp, self.CC_nodes, hip.stream_handle()),
                      "dwba_fullsplit_prep_f64")
        arr = (hip.PlainJob * n)()
        nra = (ctypes.c_int32 * n)()
        nca = (ctypes.c_int32 * n)()
        keep = []
        pos = {k: c for st in self.E for c, k in enumerate(st.members)}
        absm = [tuple(m.abs() for m in (st.ad, st.bd, st.ae, st.be)) for st in self.E]   # 4 launches per energy
        abs_plain = {k: v.abs() for k, v in self._plain_vec.items()}
        keep += [absm, abs_plain]
        for k, (ei, pi, plan, ci, cf, bf) in enumerate(self.pairs):
            st = self.E[ei]
            if k in pos:
                c = pos[k]
                av = [m[c] for m in absm[ei]]
            else:
                av = list(abs_plain[k])
            g = int(self.out_index[k])
            arr[g] = hip.PlainJob(hip.ptr(st.r_d), hip.ptr(av[0]), hip.ptr(av[1]), hip.ptr(av[2]), hip.ptr(av[3]),
                                  plan.rows, plan.cols)
            nra[g], nca[g] = plan.rows, plan.cols
        d_jobs = hip.struct_array_to_device(arr, self.dev)
        # all terms are non-negative: the O(N) form has no cancellation to lose accuracy to
        wsb = lib.dwba_il_semisep_f64_workspace_bytes(n, nra, nca, self.L_max, 0)
        ws = torch.empty(wsb, dtype=torch.uint8, device=self.dev)
        mag = torch.zeros((n, 2, self.L_max + 1), dtype=torch.float64, device=self.dev)
        hip.check(lib.dwba_il_semisep_f64(n, hip.ptr(d_jobs), nra, nca, self.L_max, 0, 0, hip.ptr(mag),
                                          hip.ptr(ws), wsb, hip.stream_handle()), "dwba_il_semisep_f64")
        return mag.cpu().numpy()[self.out_index]

    def work_flop(self):
        """Algorithmic work W = 2 N1 N2 min(D1, D2) + 2 N1 P per (energy, L,
        family), summed: exchange columns are distinct (D = P), so each pair
        contributes 2 N1_p S_p + 2 N1_p with S_p its beta support; the direct
        family with one shared beta contributes 2 R N + 2 sum_p N1_p per energy
        (R = largest N1).  Returns (W_direct, W_exchange) per launch."""
        nl = self.L_max + 1
        wd = we = 0.0
        for ei, st in enumerate(self.E):
            plans = [p[2] for p in self.pairs if p[0] == ei]
            shared = len({id(p[5]) for p in self.pairs if p[0] == ei}) == 1
            act = [p for p in plans if p.mode in (PREP_FILON, PREP_STANDARD)]
            if act:
                R = max(p.N1 for p in act)
                if shared:
                    wd += nl * (2.0 * R * st.N + 2.0 * sum(p.N1 for p in act))
                else:
                    wd += nl * sum(2.0 * p.N1 * st.N + 2.0 * p.N1 for p in act)
                we += nl * sum(2.0 * p.N1 * self._ex_support(p) + 2.0 * p.N1 for p in act)
            for p in plans:
                if p.mode == PREP_PLAIN:
                    wd += nl * (2.0 * p.N1 * p.N1 + 2.0 * p.N1)
                    we += nl * (2.0 * p.N1 * p.N1 + 2.0 * p.N1)
        return wd, we

    def block_work(self):
        """(pair grid points, MFMA flop) of one launch of the block-form MFMA
        contraction: every group sweeps n_blocks blocks of 8 points for its
        pairs; per (point, pair) the matrix cores do the four point moments
        (4 x 32 multipoles x 2 flop) and 1/8 of the block's 36 entries
        (32 x 36 x 2 flop): 544 flop."""
        if self.xb_plan is None:
            return 0, 0.0
        pp = sum(8 * g.n_blocks * g.n_pairs for g in self.xb_plan._groups_host)
        return pp, 544.0 * pp

    def tune(self, tuner=None, steps: int = 3):
        """Choose how the FP64 step runs, with the autotuner (cached per device
        and workload signature), and use it from now on; a captured graph is
        captured again.  Stage 1 times the contraction alone over every
        candidate: the wavefront-scan kernel (seg, lb), the lane kernel (lb,
        whole or split ranges at 2 or 3 waves per SIMD, job order), each with
        and without the shared-beta direct path (O(N) GEMV + matrix-core dots)
        where an energy has one direct beta.  Stage 2 times whole steps: the
        best candidate of each kind x the tail fork point x (lane kernels)
        one-stream or pipelined.  Returns stage 1's Choice (stage 2's is
        self.tuning_step), or None when the batch has no semiseparable
        contraction or DWBA_AUTOTUNE=off."""
        if self.precision != "fp64" or self.contraction != "semiseparable" or not self.plain_count:
            return None
        from . import autotune as AT
        if not AT.enabled():
            return None                     # DWBA_AUTOTUNE=off: library defaults, nothing timed
        torch, hip, lib = self.torch, self.hip, self.lib
        tuner = tuner or AT.default_tuner()
        nra, nca = self.plain_n
        Lm = self.L_max
        has_gemv = self.ss_gemv is not None
        key = AT.signature("semisep_f64", AT.device_id(self.dev), jobs=AT.bucket(self.plain_count),
                           n=AT.bucket(float(np.mean(list(nca)))), L=Lm, gemv=int(has_gemv), v=7)
        self.launch()                       # alpha/beta in place for the contraction
        torch.cuda.synchronize()
        out = torch.empty_like(self.out)
        wss = {}

        def run(c):
            seg = int(c["seg"])
            if seg not in wss:
                nb = self._semisep_ws_bytes(seg)
                wss[seg] = (torch.empty(max(nb, 1), dtype=torch.uint8, device=self.dev), nb)
            ws, nb = wss[seg]
            self.lane_order = "energy" if c.get("order", 0) else "length"
            self._contract_semisep(seg, int(c["lb"]), bool(c.get("gemv", 0)), out, ws, nb, hip.stream_handle(),
                                   bool(c.get("lanes", 0)), int(c.get("split", 0)))

        cands = AT.semisep_candidates(Lm)
        if has_gemv:
            cands = cands + [dict(c, gemv=1) for c in cands]
        choice = tuner.tune(key, cands, run, steps=steps)
        del wss
        seg, lb = int(choice.params["seg"]), int(choice.params["lb"])
        self.semisep_config = (seg, lb)
        self.semisep_lanes = bool(choice.params.get("lanes", 0))
        self.lane_split = int(choice.params.get("split", 0))
        self.lane_order = "energy" if choice.params.get("order", 0) else "length"
        self.direct_gemv = bool(choice.params.get("gemv", 0)) and has_gemv
        self.plain_ws_bytes = self._semisep_ws_bytes(seg)
        self.plain_ws = torch.empty(max(self.plain_ws_bytes, 1), dtype=torch.uint8, device=self.dev)
        self.tuning = choice
        self.tuning_step = None
        # Final choice on whole-step time: the fastest contraction of each kind
        # (shared-beta path or not, lane kernel or not, split or not, job
        # order), each with the tail chain forked before or after prep or not
        # at all, the lane kernels also pipelined (the contraction alone does
        # not show how much of it the overlapped stages hide).
        modes = (None, "start", "after_prep") if self.n_tail else (None,)
        finalists = [choice.params]
        kind = lambda q: (int(q.get("gemv", 0)), int(q.get("lanes", 0)), int(q.get("split", 0)), int(q.get("order", 0)))
        kinds = {kind(t["params"]) for t in choice.timings}
        for want in sorted(kinds):                         # (shared-beta path, lane kernel, split lanes, job order)
            same = [t for t in choice.timings if kind(t["params"]) == want]
            if same:
                best = min(same, key=lambda t: t["ms"])["params"]
                if best not in finalists:
                    finalists.append(best)
        jkey = AT.signature("scan_step", AT.device_id(self.dev), jobs=AT.bucket(self.n_pairs),
                            tails=AT.bucket(self.n_tail), n=AT.bucket(float(np.mean(list(nca)))), L=Lm,
                            c="|".join(f"{f['seg']}.{f['lb']}.{int(f.get('lanes', 0))}.{int(f.get('split', 0))}."
                                       f"{int(f.get('order', 0))}.{int(f.get('gemv', 0))}"
                                       for f in finalists), v=8)

        def apply(c):
            f = finalists[int(c["c"])]
            self.semisep_config = (int(f["seg"]), int(f["lb"]))
            self.semisep_lanes = bool(f.get("lanes", 0))
            self.lane_split = int(f.get("split", 0))
            self.lane_order = "energy" if f.get("order", 0) else "length"
            self.direct_gemv = bool(f.get("gemv", 0)) and has_gemv
            self.plain_ws_bytes = self._semisep_ws_bytes(int(f["seg"]))
            if self.plain_ws.numel() < self.plain_ws_bytes:
                self.plain_ws = torch.empty(max(self.plain_ws_bytes, 1), dtype=torch.uint8, device=self.dev)
            self.overlap = modes[int(c["overlap"])]
            self.pipeline = bool(c.get("pipe", 0))
            self.prep_grouped = bool(c.get("grp", 0))

        def run_step(c):
            apply(c)
            self.launch()

        cands = [{"c": i, "overlap": j} for i in range(len(finalists)) for j in range(len(modes))]
        if self._pipe is not None:       # the two-half pipelined step, for the lane-kernel finalists
            cands += [{"c": i, "overlap": j, "pipe": 1} for i, f in enumerate(finalists) if f.get("lanes", 0)
                      for j in range(len(modes))]
        if self.filon_groups is not None:   # Filon cell sums once per group of pairs
            cands += [{"c": i, "overlap": j, "grp": 1} for i in range(len(finalists)) for j in range(len(modes))]
        # stage 1's winner comes first; another layout must beat it by 2 %
        oc = tuner.tune(jkey, cands, run_step, steps=max(steps, 5), margin=0.02)
        apply(oc.params)
        self.tuning_overlap = oc
        self.tuning_step = {"finalists": finalists, "modes": [str(m) for m in modes], "choice": oc.params,
                            "ms": oc.ms, "timings": oc.timings}
        if self.graph is not None:
            self.graph = None
            self.capture()
        return choice

    def capture(self):
        """Capture launch() into a HIP graph (replayed by run())."""
        torch = self.torch
        s = torch.cuda.Stream(device=self.dev)
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            self.launch(s)          # warm-up outside capture
        torch.cuda.current_stream().wait_stream(s)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            self.launch(s)
        self.graph = g
        return g

    def run(self):
        if self.graph is not None:
            self.graph.replay()
        else:
            self.launch()

    def results(self):
        """(n_pairs, 2, L_max+1) numpy array in the order the pairs were given."""
        host = self.out.cpu().numpy()
        return host[self.out_index]

    def integrals(self) -> List[List[RadialDWBAIntegrals]]:
        res = self.results()
        out, k = [], 0
        counts = {}
        for (ei, pi, *_rest) in self.pairs:
            counts[ei] = counts.get(ei, 0) + 1
        for ei in range(len(self.E)):
            row = []
            for _ in range(counts.get(ei, 0)):
                row.append(RadialDWBAIntegrals({L: float(res[k, 0, L]) for L in range(self.L_max + 1)},
                                               {L: float(res[k, 1, L]) for L in range(self.L_max + 1)}))
                k += 1
            out.append(row)
        return out
