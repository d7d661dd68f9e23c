"""Drop-in replacement of the reference module ``dwba_matrix_elements``.

Keeps the names and call signatures that driver.py (:67-73, :258, :896) and
ionization.py (:94-99) import:

    radial_ME_all_L(...)      dwba_matrix_elements.py:780-1300
    radial_ME_all_L_gpu(...)  dwba_matrix_elements.py:1563-2549
    GPUCache.from_grid(grid, idx_limit=-1, max_chi_cached=20), .clear()
    RadialDWBAIntegrals, HAS_CUPY, check_cupy_runtime(), OscillatoryMethod

Both radial functions run the same MI355X path and follow the CPU semantics
of radial_ME_all_L (the reference GPU path deviates from them in several
places; the CPU values are the ones reproduced):

  host   split index, branch and phase-interval decisions (planner.py)
  GPU    dwba_fullsplit_prep_f64   alpha/beta of direct and exchange, L=0 factors
         dwba_moments_f64          M_L = sum w r^L u_f u_i
         dwba_il_semisep_f64       alpha^T K_L beta for all L in O(N) per L (default), or
         dwba_il_plain_f64         the dense O(N^2) form; K_L is never stored
                                   (set_contraction() / DWBA_CONTRACTION=dense|semiseparable)
         dwba_outer_tail_f64       Filon/Levin outer tail + analytic tail
  host   I_L = contraction + tail (+ L=0 correction), one device-to-host copy.

There is no CPU fallback: without the HIP library or a device every call
raises RuntimeError (HAS_CUPY / check_cupy_runtime() report it up front).
"""
from __future__ import annotations

import ctypes
import logging
import os
from dataclasses import dataclass, field
from typing import Dict, Literal, Optional

import numpy as np

from . import planner
from .planner import GridStats, PREP_PLAIN, PREP_FILON, PREP_STANDARD, PREP_ZERO

logger = logging.getLogger(__name__)

OscillatoryMethod = Literal["legacy", "advanced", "full_split"]

_CONTRACTIONS = ("semiseparable", "dense")
_contraction = os.environ.get("DWBA_CONTRACTION", "semiseparable")
if _contraction not in _CONTRACTIONS:
    raise ValueError(f"DWBA_CONTRACTION must be one of {_CONTRACTIONS}, got {_contraction!r}")


def set_contraction(name: str) -> None:
    """Select how alpha^T K_L beta is evaluated: "semiseparable" (O(N) per L,
    the default) or "dense" (O(N^2) per L).  Both are FP64 and agree to
    rounding; the choice is process-wide."""
    global _contraction
    if name not in _CONTRACTIONS:
        raise ValueError(f"contraction must be one of {_CONTRACTIONS}, got {name!r}")
    _contraction = name


def get_contraction() -> str:
    return _contraction

TAIL_NONE, TAIL_OUTER, TAIL_LEGACY = 0, 1, 2


def _library_present() -> bool:
    try:
        from . import _hip
        _hip.load_library()
        return True
    except Exception:
        return False


# Name kept for the drivers' GPU gate (driver.py:857-862, ionization.py:958-962):
# True when the gfx950 library loads in this process.
HAS_CUPY = _library_present()


def check_cupy_runtime() -> bool:
    """True if the HIP library initialises on the current device and runs a
    contraction (the probe of dwba_matrix_elements.py:178-193)."""
    if not HAS_CUPY:
        return False
    try:
        import torch
        from . import _hip
        _hip.ensure_init()
        r = torch.tensor([0.5, 1.0, 2.0], dtype=torch.float64, device="cuda")
        out = _hip.il_plain_f64(r, r, r, r, r, 3, 3, 0).cpu()
        return bool(torch.isfinite(out).all())
    except Exception as e:  # pragma: no cover - depends on the machine
        print(f"[Warning] GPU initialization failed: {e}")
        return False


@dataclass(frozen=True)
class RadialDWBAIntegrals:
    """I_L_direct[L], I_L_exchange[L] for L = 0..L_max (dwba_matrix_elements.py:670-685)."""
    I_L_direct: Dict[int, float]
    I_L_exchange: Dict[int, float]


@dataclass
class GPUCache:
    """Energy-level device cache (dwba_matrix_elements.py:203-614): the grid
    r, log r, Simpson weights on the device plus host prefix statistics.
    Keyed by the grid object it was built from, never by a raw address."""
    r_gpu: object
    w_gpu: object
    logr_gpu: object = None
    stats: Optional[GridStats] = None
    idx_limit: int = -1
    max_chi_cached: int = 20
    grid_ref: object = None
    extent_cache: dict = field(default_factory=dict)

    @classmethod
    def from_grid(cls, grid, idx_limit: int = -1, max_chi_cached: int = 20) -> "GPUCache":
        if not HAS_CUPY:
            raise RuntimeError("GPUCache requires the gfx950 HIP library")
        import torch
        from . import _hip
        _hip.ensure_init()
        r = np.ascontiguousarray(grid.r, dtype=np.float64)
        w = np.ascontiguousarray(grid.w_simpson, dtype=np.float64)
        wt = grid.w_trapz if hasattr(grid, "w_trapz") else np.gradient(r)
        dev = torch.device("cuda", torch.cuda.current_device())
        return cls(r_gpu=torch.as_tensor(r, device=dev), w_gpu=torch.as_tensor(w, device=dev),
                   logr_gpu=torch.as_tensor(np.log(np.maximum(r, 1e-300)), device=dev),
                   stats=GridStats.of(r, wt), idx_limit=idx_limit, max_chi_cached=max_chi_cached,
                   grid_ref=grid)

    def matches(self, grid) -> bool:
        if self.grid_ref is grid:
            return True
        r = np.asarray(grid.r)
        return (len(r) == self.stats.r.size and r[0] == self.stats.r[0] and r[-1] == self.stats.r[-1]
                and np.array_equal(r, self.stats.r))

    def clear(self) -> None:
        self.r_gpu = self.w_gpu = self.logr_gpu = None
        self.extent_cache.clear()


def _attr(obj, name, default):
    return getattr(obj, name) if hasattr(obj, name) else default


def _fit(name, arr, n):
    """Array length check of radial_ME_all_L_gpu (dwba_matrix_elements.py:1637-1672):
    longer arrays are truncated, shorter ones rejected."""
    a = np.asarray(arr, dtype=np.float64)
    if a.ndim != 1:
        raise ValueError(f"{name} must be one-dimensional")
    if len(a) < n:
        raise ValueError(f"{name} has {len(a)} pts but grid has {n}")
    return np.ascontiguousarray(a[:n])


@dataclass
class PairPlan:
    """Host decisions for one channel pair."""
    N: int
    N1: int
    L_max: int
    mode: int
    layout: planner.FilonLayout
    tail_kind: int
    oscillatory: bool
    method: str
    rows: int
    cols: int
    log1: int
    log2: int
    k_i: float
    k_f: float
    l_i: int
    l_f: int
    delta_i: float
    delta_f: float
    eta_i: float
    eta_f: float
    sigma_i: float
    sigma_f: float
    r_match: float
    r_max: float


def plan_pair(gs: GridStats, U_i, U_f, u_i, u_f, cont_i, cont_f, L_max, use_osc, method, phase_inc,
              min_grid_fraction, k_threshold, is_excitation, extent_cache=None) -> PairPlan:
    if method not in ("legacy", "advanced", "full_split"):
        raise ValueError(f"Unknown oscillatory_method: {method}")
    r = gs.r
    N = len(r)
    k_i, k_f = float(cont_i.k_au), float(cont_f.k_au)
    l_i, l_f = int(cont_i.l), int(cont_f.l)
    N1 = planner.split_index(gs, U_i, U_f, u_i, u_f, k_i, k_f, l_i, l_f,
                             int(_attr(cont_i, "idx_match", -1)), int(_attr(cont_f, "idx_match", -1)),
                             min_grid_fraction, extent_cache)
    eta_i, eta_f = float(_attr(cont_i, "eta", 0.0)), float(_attr(cont_f, "eta", 0.0))
    k_total = k_i + k_f
    oscillatory = bool(use_osc) and k_total > k_threshold
    if oscillatory:
        layout = planner.filon_layout(gs, N1, k_total, phase_inc, eta_i + eta_f)
    else:
        layout = planner.FilonLayout(PREP_PLAIN)
    mode = layout.mode
    if method == "legacy":
        tail = TAIL_LEGACY if (use_osc and N1 < N - 10 and is_excitation and L_max >= 1) else TAIL_NONE
    else:
        tail = TAIL_OUTER if (oscillatory and N1 < N - 10 and is_excitation and L_max >= 1) else TAIL_NONE
    if mode == PREP_PLAIN:
        rows, cols = N1, N1
    elif mode == PREP_ZERO:
        rows, cols = 0, N
    else:
        rows, cols = N1, N
    return PairPlan(N=N, N1=N1, L_max=int(L_max), mode=mode, layout=layout, tail_kind=tail,
                    oscillatory=oscillatory, method=method, rows=rows, cols=cols,
                    log1=int(gs.log_spaced(N1)), log2=int(gs.log_spaced(N)),
                    k_i=k_i, k_f=k_f, l_i=l_i, l_f=l_f,
                    delta_i=float(_attr(cont_i, "phase_shift", 0.0)),
                    delta_f=float(_attr(cont_f, "phase_shift", 0.0)), eta_i=eta_i, eta_f=eta_f,
                    sigma_i=float(_attr(cont_i, "sigma_l", 0.0)), sigma_f=float(_attr(cont_f, "sigma_l", 0.0)),
                    r_match=float(r[N1 - 1] if N1 > 0 else r[-1]), r_max=float(r[-1]))


def combine(plan: PairPlan, contraction: np.ndarray, scalars: np.ndarray, tails: np.ndarray) -> RadialDWBAIntegrals:
    """Assemble I_L the way radial_ME_all_L adds its pieces
    (dwba_matrix_elements.py:1112-1298): Filon contractions are zeroed when
    non-finite (oscillatory_integrals.py:2647-2649, :2773-2775), tails and the
    L = 0 correction are added in the reference's order."""
    I_dir, I_ex = {}, {}
    for L in range(plan.L_max + 1):
        d = float(contraction[0, L])
        e = float(contraction[1, L])
        if plan.oscillatory:
            d = d if np.isfinite(d) else 0.0
            e = e if np.isfinite(e) else 0.0
            if plan.tail_kind == TAIL_OUTER:
                d = d + float(tails[L])
        if L == 0:
            if abs(scalars[0]) > 1e-12:
                d += float(scalars[1]) * float(scalars[0])
            if abs(scalars[2]) > 1e-12:
                e += float(scalars[3]) * float(scalars[2])
        if plan.tail_kind == TAIL_LEGACY and L >= 1:
            d += float(tails[L])
        I_dir[L] = float(d)
        I_ex[L] = float(e)
    return RadialDWBAIntegrals(I_L_direct=I_dir, I_L_exchange=I_ex)


class _PairRunner:
    """Launches the four GPU stages of one pair on the current stream."""

    def __init__(self):
        import torch
        from . import _hip
        self.torch = torch
        self.hip = _hip
        self.lib = _hip.load_library()
        self.dev_index = _hip.ensure_init()
        self.device = torch.device("cuda", self.dev_index)
        self.levin_D, self.levin_x = planner.levin_tables(8)

    def run(self, plan: PairPlan, r_d, logr_d, w_d, u_i, u_f, chi_i, chi_f, V_diff, CC_nodes):
        torch, hip, lib = self.torch, self.hip, self.lib
        dev = self.device
        T = lambda a: torch.as_tensor(a, device=dev)
        N, L_max = plan.N, plan.L_max
        ui_d, uf_d, ci_d, cf_d, vd_d = T(u_i), T(u_f), T(chi_i), T(chi_f), T(V_diff)
        vec = torch.empty((4, N), dtype=torch.float64, device=dev)
        # packed output: [0:2(L+1)] contraction, [2(L+1):+4] scalars, then tails (L+1)
        nl = L_max + 1
        res = torch.zeros(2 * nl + 4 + nl, dtype=torch.float64, device=dev)
        scal = res[2 * nl:2 * nl + 4]
        tails = res[2 * nl + 4:]
        lay = plan.layout
        nodes_d = None
        par_d = None
        if plan.mode == PREP_FILON and lay.coulomb:
            nodes_d = torch.empty(10002, dtype=torch.float64, device=dev)
            par_d = T(np.array([lay.p_start, lay.p_end, lay.k_total, lay.eta_total], dtype=np.float64))
        job = (hip.PrepJob * 1)()
        job[0] = hip.PrepJob(hip.ptr(r_d), hip.ptr(logr_d), hip.ptr(w_d), hip.ptr(ui_d), hip.ptr(uf_d),
                             hip.ptr(ci_d), hip.ptr(cf_d), hip.ptr(vd_d), hip.ptr(nodes_d),
                             hip.ptr(vec[0]), hip.ptr(vec[1]), hip.ptr(vec[2]), hip.ptr(vec[3]), hip.ptr(scal),
                             lay.p_start, lay.p_step, lay.p_end, N, plan.N1, lay.n_pnodes, plan.mode,
                             plan.log1, plan.log2)
        job_d = hip.struct_array_to_device(job, dev)
        cx, cxa = hip.f64_array(planner.cc_reference(CC_nodes)[0])
        cw, cwa = hip.f64_array(planner.cc_reference(CC_nodes)[1])
        sh = hip.stream_handle()
        hip.check(lib.dwba_fullsplit_prep_f64(1, hip.ptr(job_d), hip.ptr(par_d), int(par_d is not None),
                                              cx, cw, int(CC_nodes), sh), "dwba_fullsplit_prep_f64")
        keep = [job_d, nodes_d, par_d, cxa, cwa]
        if plan.tail_kind != TAIL_NONE:
            mom = torch.empty(nl, dtype=torch.float64, device=dev)
            ptrs = T(np.array([hip.ptr(r_d), hip.ptr(w_d), hip.ptr(ui_d), hip.ptr(uf_d)], dtype=np.int64))
            n_d = T(np.array([N], dtype=np.int32))
            hip.check(lib.dwba_moments_f64(1, hip.ptr(ptrs[0:1]), hip.ptr(ptrs[1:2]), hip.ptr(ptrs[2:3]),
                                           hip.ptr(ptrs[3:4]), hip.ptr(n_d), L_max, hip.ptr(mom), sh),
                      "dwba_moments_f64")
            tj = (hip.TailJob * 1)()
            r_m = plan.r_match if plan.tail_kind == TAIL_LEGACY else plan.r_match
            tj[0] = hip.TailJob(plan.k_i, plan.delta_i, plan.eta_i, plan.sigma_i, plan.k_f, plan.delta_f,
                                plan.eta_f, plan.sigma_f, plan.l_i, plan.l_f, r_m, plan.r_max,
                                hip.ptr(mom), hip.ptr(tails), L_max, plan.tail_kind)
            tj_d = hip.struct_array_to_device(tj, dev)
            seg, dmax = planner.filon_limits()
            LD, LDa = hip.f64_array(self.levin_D.ravel())
            LX, LXa = hip.f64_array(self.levin_x)
            twb = lib.dwba_outer_tail_workspace_bytes(1, L_max)
            tws = torch.empty(twb, dtype=torch.uint8, device=dev)
            hip.check(lib.dwba_outer_tail_f64(1, hip.ptr(tj_d), L_max, seg, dmax, LD, LX, hip.ptr(tws), twb, sh),
                      "dwba_outer_tail_f64")
            keep += [mom, ptrs, n_d, tj_d, LDa, LXa, tws]
        if plan.rows > 0:
            out = res[:2 * nl].view(2, nl)
            if _contraction == "semiseparable":
                pj = (hip.PlainJob * 1)()
                pj[0] = hip.PlainJob(hip.ptr(r_d), hip.ptr(vec[0]), hip.ptr(vec[1]), hip.ptr(vec[2]),
                                     hip.ptr(vec[3]), plan.rows, plan.cols)
                pj_d = hip.struct_array_to_device(pj, dev)
                nr = (ctypes.c_int32 * 1)(plan.rows)
                nc = (ctypes.c_int32 * 1)(plan.cols)
                ws_bytes = lib.dwba_il_semisep_f64_workspace_bytes(1, nr, nc, L_max, 0)
                ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
                hip.check(lib.dwba_il_semisep_f64(1, hip.ptr(pj_d), nr, nc, L_max, 0, 0, hip.ptr(out), hip.ptr(ws),
                                                  ws_bytes, sh), "dwba_il_semisep_f64")
                keep += [pj_d, ws]
            else:
                ws_bytes = lib.dwba_il_plain_f64_workspace_bytes(plan.rows, plan.cols, L_max)
                ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
                hip.check(lib.dwba_il_plain_f64(hip.ptr(r_d), plan.rows, plan.cols, hip.ptr(vec[0]),
                                                hip.ptr(vec[1]), hip.ptr(vec[2]), hip.ptr(vec[3]), L_max,
                                                hip.ptr(out), hip.ptr(ws), ws_bytes, sh), "dwba_il_plain_f64")
                keep.append(ws)
        host = res.cpu().numpy()          # the one device-to-host copy (synchronises)
        del keep
        return (host[:2 * nl].reshape(2, nl), host[2 * nl:2 * nl + 4], host[2 * nl + 4:])


_runner: Optional[_PairRunner] = None


def _get_runner() -> _PairRunner:
    global _runner
    if not HAS_CUPY:
        raise RuntimeError("radial_ME_all_L_gpu: the gfx950 HIP library is not available "
                           "(build it with __graft_entry__.build()); there is no CPU fallback")
    if _runner is None:
        _runner = _PairRunner()
    return _runner


def _radial(grid, V_core_array, U_i_array, bound_i, bound_f, cont_i, cont_f, L_max,
            use_oscillatory_quadrature, oscillatory_method, CC_nodes, phase_increment,
            min_grid_fraction, k_threshold, gpu_cache, U_f_array) -> RadialDWBAIntegrals:
    if L_max < 0:
        raise ValueError("radial_ME_all_L: L_max must be >= 0.")
    runner = _get_runner()
    torch = runner.torch
    r = np.ascontiguousarray(grid.r, dtype=np.float64)
    N = len(r)
    if gpu_cache is not None and gpu_cache.stats is not None and gpu_cache.r_gpu is not None \
            and gpu_cache.matches(grid):
        gs = gpu_cache.stats
        r_d, w_d, logr_d = gpu_cache.r_gpu, gpu_cache.w_gpu, gpu_cache.logr_gpu
        ext = gpu_cache.extent_cache
    else:
        wt = grid.w_trapz if hasattr(grid, "w_trapz") else np.gradient(r)
        gs = GridStats.of(r, wt)
        dev = runner.device
        r_d = torch.as_tensor(r, device=dev)
        w_d = torch.as_tensor(np.ascontiguousarray(grid.w_simpson, dtype=np.float64), device=dev)
        logr_d = torch.as_tensor(np.log(np.maximum(r, 1e-300)), device=dev)
        ext = None
    V_core = _fit("V_core", V_core_array, N)
    U_i = _fit("U_i", U_i_array, N)
    U_f = _fit("U_f", U_f_array, N) if U_f_array is not None else None
    chi_i = _fit("cont_i.chi", cont_i.chi_of_r, N)
    chi_f = _fit("cont_f.chi", cont_f.chi_of_r, N)
    u_i = _fit("bound_i.u", bound_i.u_of_r, N)
    u_f = _fit("bound_f.u", bound_f.u_of_r, N)
    is_exc = not hasattr(bound_f, "k_au")
    plan = plan_pair(gs, U_i, U_f, u_i, u_f, cont_i, cont_f, L_max, use_oscillatory_quadrature,
                     oscillatory_method, phase_increment, min_grid_fraction, k_threshold, is_exc, ext)
    contraction, scalars, tails = runner.run(plan, r_d, logr_d, w_d, u_i, u_f, chi_i, chi_f,
                                             V_core - U_i, CC_nodes)
    return combine(plan, contraction, scalars, tails)


def radial_ME_all_L(grid, V_core_array, U_i_array, bound_i, bound_f, cont_i, cont_f, L_max: int,
                    use_oscillatory_quadrature: bool = True, oscillatory_method: OscillatoryMethod = "advanced",
                    CC_nodes: int = 5, phase_increment: float = 1.5708, min_grid_fraction: float = 0.10,
                    k_threshold: float = 0.5, U_f_array: Optional[np.ndarray] = None) -> RadialDWBAIntegrals:
    """Direct and exchange I_L, L = 0..L_max (signature of
    dwba_matrix_elements.py:780-799), computed on the MI355X."""
    return _radial(grid, V_core_array, U_i_array, bound_i, bound_f, cont_i, cont_f, L_max,
                   use_oscillatory_quadrature, oscillatory_method, CC_nodes, phase_increment,
                   min_grid_fraction, k_threshold, None, U_f_array)


def radial_ME_all_L_gpu(grid, V_core_array, U_i_array, bound_i, bound_f, cont_i, cont_f, L_max: int,
                        use_oscillatory_quadrature: bool = True,
                        oscillatory_method: OscillatoryMethod = "advanced", CC_nodes: int = 5,
                        phase_increment: float = 1.5708, gpu_block_size: int = 0,
                        min_grid_fraction: float = 0.10, k_threshold: float = 0.5,
                        gpu_memory_mode: str = "auto", gpu_memory_threshold: float = 0.8,
                        gpu_cache: Optional[GPUCache] = None,
                        U_f_array: Optional[np.ndarray] = None) -> RadialDWBAIntegrals:
    """Signature of dwba_matrix_elements.py:1563-1584.  gpu_block_size,
    gpu_memory_mode and gpu_memory_threshold are accepted and ignored: K_L is
    generated on chip, so there is no memory tier to choose."""
    return _radial(grid, V_core_array, U_i_array, bound_i, bound_f, cont_i, cont_f, L_max,
                   use_oscillatory_quadrature, oscillatory_method, CC_nodes, phase_increment,
                   min_grid_fraction, k_threshold, gpu_cache, U_f_array)
