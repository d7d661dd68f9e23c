"""ctypes binding of the gfx950 library (include/dwba_hip.h).

torch is imported before the library is loaded so that the HIP runtime torch
ships (soname libamdhip64.so.7) is the one the library binds to: device
pointers and streams created by torch are then valid in the library.

There is no fallback: if the library or a GPU is missing, every entry point
raises RuntimeError.
"""
from __future__ import annotations

import ctypes
import os
import threading

import torch  # noqa: F401  (must precede the CDLL load, see module docstring)

from ._build import lib_path

DWBA_OK = 0
DWBA_ERR_ARG = 1
DWBA_ERR_HIP = 2
DWBA_ERR_NOT_INIT = 3
DWBA_ERR_WORKSPACE = 4

_lock = threading.Lock()
_lib = None
_init_dev = None

c_i32 = ctypes.c_int32
c_sz = ctypes.c_size_t
c_vp = ctypes.c_void_p
c_i32p = ctypes.POINTER(ctypes.c_int32)


class PlainJob(ctypes.Structure):
    """dwba_plain_job of include/dwba_hip.h."""
    _fields_ = [("r", c_vp), ("alpha_dir", c_vp), ("beta_dir", c_vp), ("alpha_ex", c_vp),
                ("beta_ex", c_vp), ("n_rows", c_i32), ("n_cols", c_i32)]


class PrepJob(ctypes.Structure):
    """dwba_prep_job of include/dwba_hip.h."""
    _fields_ = [("r", c_vp), ("logr", c_vp), ("w", c_vp), ("u_i", c_vp), ("u_f", c_vp), ("chi_i", c_vp),
                ("chi_f", c_vp), ("V_diff", c_vp), ("phase_nodes", c_vp), ("alpha_dir", c_vp),
                ("beta_dir", c_vp), ("alpha_ex", c_vp), ("beta_ex", c_vp), ("scalars", c_vp),
                ("p_start", ctypes.c_double), ("p_step", ctypes.c_double), ("p_end", ctypes.c_double),
                ("N", c_i32), ("N1", c_i32), ("n_pnodes", c_i32), ("mode", c_i32), ("log1", c_i32),
                ("log2", c_i32)]


class AmpPair(ctypes.Structure):
    """dwba_amp_pair of include/dwba_hip.h."""
    _fields_ = [("row", c_i32), ("type", c_i32), ("l_f", c_i32), ("pad_", c_i32), ("scale", ctypes.c_double)]


class CwWave(ctypes.Structure):
    """dwba_cw_wave of include/dwba_hip.h."""
    _fields_ = [("r", c_vp), ("U", c_vp), ("k", ctypes.c_double), ("r1_ratio", ctypes.c_double),
                ("chi0", ctypes.c_double), ("chi1", ctypes.c_double), ("n", c_i32), ("l", c_i32), ("start", c_i32),
                ("method", c_i32), ("renorm", c_i32), ("pad_", c_i32)]


class CwFinish(ctypes.Structure):
    """dwba_cw_finish of include/dwba_hip.h."""
    _fields_ = [("r", c_vp), ("w_trapz", c_vp), ("k", ctypes.c_double), ("delta", ctypes.c_double),
                ("eta", ctypes.c_double), ("sigma", ctypes.c_double), ("n", c_i32), ("l", c_i32),
                ("idx_match", c_i32), ("tail", c_i32), ("coulomb", c_i32), ("pad_", c_i32)]


class CwBypass(ctypes.Structure):
    """dwba_cw_bypass of include/dwba_hip.h."""
    _fields_ = [("r", c_vp), ("w_trapz", c_vp), ("U", c_vp), ("k", ctypes.c_double), ("n", c_i32), ("l", c_i32)]


class TailJob(ctypes.Structure):
    """dwba_tail_job of include/dwba_hip.h."""
    _fields_ = [("k_a", ctypes.c_double), ("delta_a", ctypes.c_double), ("eta_a", ctypes.c_double),
                ("sigma_a", ctypes.c_double), ("k_b", ctypes.c_double), ("delta_b", ctypes.c_double),
                ("eta_b", ctypes.c_double), ("sigma_b", ctypes.c_double), ("l_a", c_i32), ("l_b", c_i32),
                ("r_m", ctypes.c_double), ("r_max", ctypes.c_double), ("moments", c_vp), ("out", c_vp),
                ("L_max", c_i32), ("kind", c_i32)]


class ScanEnergy(ctypes.Structure):
    """dwba_scan_energy of include/dwba_hip.h (host pointers)."""
    _fields_ = [("r", c_vp), ("w_simpson", c_vp), ("w_trapz", c_vp), ("V_core", c_vp), ("U_i", c_vp),
                ("U_f", c_vp), ("u_i", c_vp), ("n", c_i32), ("pad_", c_i32)]


class ScanWave(ctypes.Structure):
    """dwba_scan_wave of include/dwba_hip.h (host pointer)."""
    _fields_ = [("chi", c_vp), ("k", ctypes.c_double), ("phase_shift", ctypes.c_double),
                ("eta", ctypes.c_double), ("sigma_l", ctypes.c_double), ("l", c_i32), ("idx_match", c_i32),
                ("n", c_i32), ("pad_", c_i32)]


class ScanPair(ctypes.Structure):
    """dwba_scan_pair of include/dwba_hip.h (host pointer)."""
    _fields_ = [("u_f", c_vp), ("n_u_f", c_i32), ("energy", c_i32), ("wave_i", c_i32), ("wave_f", c_i32),
                ("excitation", c_i32), ("pad_", c_i32)]


class ScanOptions(ctypes.Structure):
    """dwba_scan_options of include/dwba_hip.h."""
    _fields_ = [("L_max", c_i32), ("use_oscillatory_quadrature", c_i32), ("method", c_i32), ("CC_nodes", c_i32),
                ("phase_increment", ctypes.c_double), ("min_grid_fraction", ctypes.c_double),
                ("k_threshold", ctypes.c_double)]


class ScanPairPlan(ctypes.Structure):
    """dwba_scan_pair_plan of include/dwba_hip.h."""
    _fields_ = [("p_start", ctypes.c_double), ("p_step", ctypes.c_double), ("p_end", ctypes.c_double),
                ("r_match", ctypes.c_double), ("N1", c_i32), ("mode", c_i32), ("tail", c_i32), ("rows", c_i32),
                ("cols", c_i32), ("log1", c_i32), ("log2", c_i32), ("n_pnodes", c_i32), ("coulomb", c_i32),
                ("oscillatory", c_i32)]


SCAN_METHODS = {"legacy": 0, "advanced": 1, "full_split": 2}   # DWBA_METHOD_* of include/dwba_hip.h

c_dp = ctypes.POINTER(ctypes.c_double)

# name -> (restype, argtypes)
_SIGS = {
    "dwba_version": (c_i32, []),
    "dwba_init": (c_i32, [c_i32]),
    "dwba_shutdown": (c_i32, []),
    "dwba_last_error": (ctypes.c_char_p, []),
    "dwba_il_plain_f64_workspace_bytes": (c_sz, [c_i32, c_i32, c_i32]),
    "dwba_il_plain_f64": (c_i32, [c_vp, c_i32, c_i32, c_vp, c_vp, c_vp, c_vp, c_i32, c_vp, c_vp, c_sz, c_vp]),
    "dwba_il_plain_f64_noalloc": (c_i32, [c_vp, c_i32, c_i32, c_vp, c_vp, c_vp, c_vp, c_i32, c_vp, c_vp]),
    "dwba_il_plain_f64_batch_workspace_bytes": (c_sz, [c_i32, c_i32p, c_i32p, c_i32]),
    "dwba_il_plain_f64_batch": (c_i32, [c_i32, c_vp, c_i32p, c_i32p, c_i32, c_vp, c_vp, c_sz, c_vp]),
    "dwba_il_semisep_f64_workspace_bytes": (c_sz, [c_i32, c_i32p, c_i32p, c_i32, c_i32]),
    "dwba_il_semisep_f64": (c_i32, [c_i32, c_vp, c_i32p, c_i32p, c_i32, c_i32, c_i32, c_vp, c_vp, c_sz, c_vp]),
    "dwba_fullsplit_prep_f64": (c_i32, [c_i32, c_vp, c_vp, c_i32, c_dp, c_dp, c_i32, c_vp]),
    "dwba_moments_f64": (c_i32, [c_i32, c_vp, c_vp, c_vp, c_vp, c_vp, c_i32, c_vp, c_vp]),
    "dwba_outer_tail_workspace_bytes": (c_sz, [c_i32, c_i32]),
    "dwba_outer_tail_f64": (c_i32, [c_i32, c_vp, c_i32, c_i32, ctypes.c_double, c_dp, c_dp, c_vp, c_sz, c_vp]),
    "dwba_cast_f64_f32": (c_i32, [c_vp, c_vp, ctypes.c_int64, c_vp]),
    "dwba_xgemm_partial_doubles": (c_sz, [c_i32, c_i32, c_i32]),
    "dwba_xgemm_f32": (c_i32, [c_i32, c_vp, c_i32, c_vp, c_i32, c_vp, c_vp, c_i32, c_vp, c_i32, c_vp]),
    "dwba_xgemm_bf16x3": (c_i32, [c_i32, c_vp, c_i32, c_vp, c_i32, c_vp, c_vp, c_i32, c_vp, c_i32, c_vp]),
    "dwba_split_bf16": (c_i32, [c_vp, c_vp, c_vp, ctypes.c_int64, c_vp]),
    "dwba_xblock_table_bytes": (c_sz, [c_i32, c_i32]),
    "dwba_xblock_partial_doubles": (c_sz, [c_i32]),
    "dwba_xblock_tables": (c_i32, [c_i32, c_vp, c_i32, c_i32, c_vp]),
    "dwba_xblock_contract": (c_i32, [c_i32, c_i32, c_vp, c_i32, c_vp, c_i32, c_vp, c_vp, c_vp]),
    "dwba_dgemv_f64": (c_i32, [c_i32, c_vp, c_i32, c_i32, c_vp]),
    "dwba_semisep_gemv_f64": (c_i32, [c_i32, c_vp, c_i32, c_vp]),
    "dwba_il_semisep_lanes_f64": (c_i32, [c_i32, c_vp, c_i32, c_i32, c_vp, c_vp]),
    "dwba_ddot_f64": (c_i32, [c_i32, c_vp, c_i32, c_vp, c_vp]),
    "dwba_ddot_tiled_f64": (c_i32, [c_i32, c_vp, c_vp, c_i32, c_vp, c_vp]),
    "dwba_ddot_mfma_f64": (c_i32, [c_i32, c_vp, c_vp, c_i32, c_vp, c_vp]),
    "dwba_ddot_stage_f64": (c_i32, [c_i32, c_vp, c_vp, c_i32, c_vp, c_vp]),
    "dwba_semisep_lane_table_stride": (c_i32, [c_i32, c_i32]),
    "dwba_filon_cells_f64": (c_i32, [c_i32, c_vp, c_vp, c_vp, c_vp, c_i32, c_vp]),
    "dwba_filon_fused_f64": (c_i32, [c_i32, c_vp, c_vp, c_vp, c_vp]),
    "dwba_semisep_lane_tables_f64": (c_i32, [c_i32, c_vp, c_i32, c_i32, c_i32, c_vp]),
    "dwba_il_semisep_lanes_tab_f64": (c_i32, [c_i32, c_vp, c_vp, c_vp, c_i32, c_i32, c_vp, c_vp]),
    "dwba_il_semisep_lanes_split_f64": (c_i32, [c_i32, c_vp, c_vp, c_vp, c_i32, c_i32, c_i32, c_vp, c_vp]),
    "dwba_semisep_lane_stage_group": (c_i32, [c_i32, c_i32]),
    "dwba_il_semisep_lanes_stage_f64": (c_i32, [c_i32, c_vp, c_vp, c_vp, c_i32, c_i32, c_vp, c_vp]),
    "dwba_batch_finalize": (c_i32, [c_i32, c_vp, c_i32, c_vp, c_vp]),
    "dwba_il_backward_f64": (c_i32, [c_vp, c_vp, c_i32, c_i32, c_i32, c_vp, c_vp, c_vp, c_i32, c_vp, c_vp, c_vp]),
    "dwba_il_backward_f32": (c_i32, [c_vp, c_vp, c_i32, c_i32, c_i32, c_vp, c_vp, c_vp, c_i32, c_vp, c_vp, c_vp]),
    "dwba_il_backward_semisep_f64_workspace_bytes": (c_sz, [c_i32, c_i32, c_i32]),
    "dwba_il_backward_semisep_f64": (c_i32, [c_vp, c_vp, c_i32, c_i32, c_i32, c_vp, c_vp, c_vp, c_i32, c_vp, c_vp,
                                             c_vp, c_sz, c_vp]),
    "dwba_weight_rows": (c_i32, [c_vp, c_i32, c_i32, c_vp, c_i32, c_vp, c_i32, c_vp]),
    "dwba_ylm_f64": (c_i32, [c_i32, c_i32, c_i32, c_vp, c_vp, c_vp]),
    "dwba_amp_coeffs_f64": (c_i32, [c_i32, c_vp, c_i32, c_i32, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "dwba_amp_sum_f64": (c_i32, [c_i32, c_vp, c_vp, c_i32, c_i32p, c_i32p, c_i32, c_i32, c_i32, c_vp, c_vp, c_vp,
                                 c_vp, c_vp]),
    "dwba_dcs_f64": (c_i32, [c_i32, c_i32, c_i32, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "dwba_sphere_trapz_f64": (c_i32, [c_i32, c_i32, c_vp, c_vp, c_vp, c_vp]),
    "dwba_pw_cumulate_f64": (c_i32, [c_i32, c_vp, c_i32, c_i32, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "dwba_ion_spin_terms_f64": (c_i32, [c_i32, c_vp, c_vp, c_i32, c_vp, c_vp, c_vp, c_vp]),
    "dwba_slot_sums_f64": (c_i32, [c_i32, c_vp, c_vp, c_i32, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "dwba_cw_workspace_bytes": (c_sz, [c_i32, c_i32]),
    "dwba_cw_propagate_f64": (c_i32, [c_i32, c_vp, c_i32, c_vp, c_vp, c_vp, c_sz, c_vp, c_vp]),
    "dwba_cw_tail_sums_f64": (c_i32, [c_i32, c_vp, c_i32, c_vp, c_vp, c_vp]),
    "dwba_cw_stitch_f64": (c_i32, [c_i32, c_vp, c_i32, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "dwba_cw_bypass_f64": (c_i32, [c_i32, c_vp, c_i32, c_vp, c_vp, c_vp]),
    "dwba_scan_plan": (c_i32, [c_i32, c_vp, c_i32, c_vp, c_i32, c_vp, c_vp, c_vp]),
    "dwba_scan_plan_pairs": (c_i32, [c_i32, c_vp, c_i32, c_vp, c_i32, c_vp, c_vp, c_vp]),
    "dwba_scan_launch": (c_i32, [c_vp, c_vp]),
    "dwba_scan_capture": (c_i32, [c_vp, c_vp]),
    "dwba_scan_replay": (c_i32, [c_vp, c_vp]),
    "dwba_scan_results": (c_vp, [c_vp]),
    "dwba_scan_read_results": (c_i32, [c_vp, c_vp, c_vp]),
    "dwba_scan_info": (c_i32, [c_vp, c_vp, c_i32]),
    "dwba_scan_destroy": (c_i32, [c_vp]),
}


def struct_array_to_device(structs, device):
    """Copy a ctypes array of structures into a device uint8 tensor."""
    raw = ctypes.string_at(ctypes.addressof(structs), ctypes.sizeof(structs))
    return torch.frombuffer(bytearray(raw), dtype=torch.uint8).to(device)


def f64_array(x):
    """Contiguous host float64 ctypes view (kept alive by the returned tuple)."""
    import numpy as np
    a = np.ascontiguousarray(x, dtype=np.float64)
    return a.ctypes.data_as(c_dp), a


def exported_symbols():
    """Names of every C-ABI function this binding expects the library to export."""
    return sorted(_SIGS)


def load_library(path: str | None = None):
    """Load the shared library (no GPU is touched)."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        p = path or lib_path()
        if not os.path.exists(p):
            raise RuntimeError(f"gfx950 library not built: {p} is missing (run __graft_entry__.build())")
        lib = ctypes.CDLL(p)
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return lib


def last_error() -> str:
    return load_library().dwba_last_error().decode(errors="replace")


def check(code: int, what: str):
    if code == DWBA_OK:
        return
    msg = f"{what}: {last_error()}"
    if code in (DWBA_ERR_ARG, DWBA_ERR_WORKSPACE):
        raise ValueError(msg)
    raise RuntimeError(msg)


def device_available() -> bool:
    try:
        return torch.cuda.is_available() and torch.cuda.device_count() > 0
    except Exception:
        return False


def ensure_init(device: int | None = None) -> int:
    """Initialise the library on `device` (default: torch's current device)."""
    global _init_dev
    lib = load_library()
    if not device_available():
        raise RuntimeError("no HIP device visible: the gfx950 path has no CPU fallback")
    dev = torch.cuda.current_device() if device is None else int(device)
    if _init_dev != dev:
        check(lib.dwba_init(dev), "dwba_init")
        _init_dev = dev
    return dev


def stream_handle(stream=None) -> int:
    s = torch.cuda.current_stream() if stream is None else stream
    return int(s.cuda_stream)


def ptr(t) -> int:
    return int(t.data_ptr()) if t is not None else 0


def _require_f64_cuda(name, t, n):
    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
        raise ValueError(f"{name} must be a contiguous float64 CUDA tensor")
    if t.numel() < n:
        raise ValueError(f"{name} has {t.numel()} elements, need >= {n}")


def il_plain_f64(r, alpha_dir, beta_dir, alpha_ex, beta_ex, n_rows: int, n_cols: int, L_max: int,
                 out=None, stream=None):
    """dwba_il_plain_f64 on torch tensors.  Returns a (2, L_max+1) float64 CUDA
    tensor: row 0 direct, row 1 exchange."""
    ensure_init()
    lib = load_library()
    for nm, t, n in (("r", r, n_cols), ("alpha_dir", alpha_dir, n_rows), ("beta_dir", beta_dir, n_cols),
                     ("alpha_ex", alpha_ex, n_rows), ("beta_ex", beta_ex, n_cols)):
        _require_f64_cuda(nm, t, n)
    ws_bytes = lib.dwba_il_plain_f64_workspace_bytes(n_rows, n_cols, L_max)
    if ws_bytes == 0:
        raise ValueError(f"il_plain_f64: invalid shape n_rows={n_rows} n_cols={n_cols} L_max={L_max}")
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=r.device)
    if out is None:
        out = torch.empty((2, L_max + 1), dtype=torch.float64, device=r.device)
    check(lib.dwba_il_plain_f64(ptr(r), n_rows, n_cols, ptr(alpha_dir), ptr(beta_dir), ptr(alpha_ex),
                                ptr(beta_ex), L_max, ptr(out), ptr(ws), ws_bytes, stream_handle(stream)),
          "dwba_il_plain_f64")
    out._dwba_ws = ws  # keep the workspace alive until the tensor is released
    return out


def il_semisep_f64_batch(jobs, L_max: int, seg: int = 0, stream=None, lb: int = 0, lanes: bool = False,
                         tables: bool = False, split: bool = False, waves_per_simd: int = 0):
    """dwba_il_semisep_f64 (or, with lanes, dwba_il_semisep_lanes_f64) on the
    same job tuples as il_plain_f64_batch (n_rows <= n_cols).  With lanes and
    tables: one grid table per job (dwba_semisep_lane_tables_f64) and the
    table-driven kernel (split: the two-wavefront form), the jobs passed in
    reverse order with d_out_index mapping them back.  Returns a
    (n_jobs, 2, L_max+1) float64 CUDA tensor."""
    ensure_init()
    lib = load_library()
    n = len(jobs)
    if n == 0:
        raise ValueError("il_semisep_f64_batch: no jobs")
    dev = jobs[0][0].device
    arr = (PlainJob * n)()
    nr = (ctypes.c_int32 * n)()
    nc = (ctypes.c_int32 * n)()
    for k, (r, ad, bd, ae, be, n_rows, n_cols) in enumerate(jobs):
        for nm, t, m in (("r", r, n_cols), ("alpha_dir", ad, n_rows), ("beta_dir", bd, n_cols),
                         ("alpha_ex", ae, n_rows), ("beta_ex", be, n_cols)):
            _require_f64_cuda(f"job {k} {nm}", t, m)
        arr[k] = PlainJob(ptr(r), ptr(ad), ptr(bd), ptr(ae), ptr(be), int(n_rows), int(n_cols))
        nr[k] = int(n_rows)
        nc[k] = int(n_cols)
    d_jobs = struct_array_to_device(arr, dev)
    ws_bytes = lib.dwba_il_semisep_f64_workspace_bytes(n, nr, nc, L_max, seg)
    if ws_bytes == 0:
        raise ValueError("il_semisep_f64_batch: invalid shapes (need 0 <= n_rows <= n_cols, 0 <= L_max <= 63)")
    out = torch.empty((n, 2, L_max + 1), dtype=torch.float64, device=dev)
    if lanes and tables:
        import numpy as np
        from .batch import LaneGrid
        lbv = lb or 8
        stride = lib.dwba_semisep_lane_table_stride(L_max, lbv)
        if stride <= 0:
            raise ValueError(f"il_semisep_f64_batch: bad (L_max, lb) = ({L_max}, {lb})")
        tabs = [torch.empty(max(int(nc[k]), 1) * stride, dtype=torch.float64, device=dev) for k in range(n)]
        live = [k for k in range(n) if nc[k] > 0]
        grids = (LaneGrid * max(len(live), 1))()
        for g, k in enumerate(live):
            grids[g] = LaneGrid(ptr(jobs[k][0]), ptr(tabs[k]), int(nc[k]), 0)
        keep = [tabs]
        if live:
            d_grids = struct_array_to_device(grids, dev)
            keep.append(d_grids)
            check(lib.dwba_semisep_lane_tables_f64(len(live), ptr(d_grids), max(int(nc[k]) for k in live), L_max,
                                                   lbv, stream_handle(stream)), "dwba_semisep_lane_tables_f64")
        rev = (PlainJob * n)()
        for p in range(n):
            rev[p] = arr[n - 1 - p]
        d_rev = struct_array_to_device(rev, dev)
        d_tabs = torch.as_tensor(np.asarray([ptr(tabs[n - 1 - p]) for p in range(n)], dtype=np.int64), device=dev)
        d_oidx = torch.as_tensor(np.arange(n - 1, -1, -1, dtype=np.int32), device=dev)
        if split:
            check(lib.dwba_il_semisep_lanes_split_f64(n, ptr(d_rev), ptr(d_tabs), ptr(d_oidx), L_max, lbv,
                                                      waves_per_simd, ptr(out), stream_handle(stream)),
                  "dwba_il_semisep_lanes_split_f64")
        else:
            check(lib.dwba_il_semisep_lanes_tab_f64(n, ptr(d_rev), ptr(d_tabs), ptr(d_oidx), L_max, lbv, ptr(out),
                                                    stream_handle(stream)), "dwba_il_semisep_lanes_tab_f64")
        out._dwba_keep = (d_jobs, d_rev, d_tabs, d_oidx, keep)
        return out
    if lanes:
        check(lib.dwba_il_semisep_lanes_f64(n, ptr(d_jobs), L_max, lb, ptr(out), stream_handle(stream)),
              "dwba_il_semisep_lanes_f64")
        out._dwba_keep = (d_jobs,)
        return out
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
    check(lib.dwba_il_semisep_f64(n, ptr(d_jobs), nr, nc, L_max, seg, lb, ptr(out), ptr(ws), ws_bytes,
                                  stream_handle(stream)), "dwba_il_semisep_f64")
    out._dwba_keep = (ws, d_jobs)
    return out


def il_semisep_stage_f64_batch(jobs, L_max: int, stream=None, pad: bool = True, lb: int = 0):
    """dwba_il_semisep_lanes_stage_f64 on the job tuples of il_semisep_f64_batch
    (lb = multipoles per lane, 4 or 8; 0 = 4).  Jobs on the same grid tensor
    share one grid table; pad: each grid's jobs (in reverse order) are padded
    to whole groups of dwba_semisep_lane_stage_group(L_max, lb) with empty jobs that d_out_index
    marks unwritten.  pad=False passes the jobs as given (no padding), so a
    group whose streams are on different grids reads back NaN.  Returns a
    (n_jobs, 2, L_max+1) float64 CUDA tensor."""
    import numpy as np
    from .batch import LaneGrid
    ensure_init()
    lib = load_library()
    n = len(jobs)
    if n == 0:
        raise ValueError("il_semisep_stage_f64_batch: no jobs")
    lbv = lb or 4
    G = int(lib.dwba_semisep_lane_stage_group(L_max, lbv))
    if G <= 0:
        raise ValueError(f"il_semisep_stage_f64_batch: (L_max, lb) = ({L_max}, {lbv}) not supported")
    dev = jobs[0][0].device
    arr = []
    for k, (r, ad, bd, ae, be, n_rows, n_cols) in enumerate(jobs):
        for nm, t, m in (("r", r, n_cols), ("alpha_dir", ad, n_rows), ("beta_dir", bd, n_cols),
                         ("alpha_ex", ae, n_rows), ("beta_ex", be, n_cols)):
            _require_f64_cuda(f"job {k} {nm}", t, m)
        arr.append(PlainJob(ptr(r), ptr(ad), ptr(bd), ptr(ae), ptr(be), int(n_rows), int(n_cols)))
    stride = lib.dwba_semisep_lane_table_stride(L_max, lbv)
    # one table per grid tensor, covering the longest job on it
    groups = {}
    for k in range(n):
        groups.setdefault(jobs[k][0].data_ptr(), []).append(k)
    tabs, gl, tab_of = [], [], {}
    for g, ks in groups.items():
        npts = max(int(jobs[k][6]) for k in ks)
        t = torch.empty(max(npts, 1) * stride, dtype=torch.float64, device=dev)
        tabs.append(t)
        tab_of[g] = ptr(t)
        if npts > 0:
            gl.append(LaneGrid(ptr(jobs[ks[0]][0]), ptr(t), npts, 0))
    if gl:
        d_grids = struct_array_to_device((LaneGrid * len(gl))(*gl), dev)
        check(lib.dwba_semisep_lane_tables_f64(len(gl), ptr(d_grids), max(g.n for g in gl), L_max, lbv,
                                               stream_handle(stream)), "dwba_semisep_lane_tables_f64")
    else:
        d_grids = None
    order, tp = [], []
    if pad:
        for g, ks in groups.items():
            body = ks[::-1] + [-1] * ((-len(ks)) % G)
            order += body
            tp += [tab_of[g]] * len(body)
    else:
        order = list(range(n))
        tp = [tab_of[jobs[k][0].data_ptr()] for k in order]
    sj = (PlainJob * len(order))()
    for p, k in enumerate(order):
        sj[p] = arr[k] if k >= 0 else PlainJob(arr[0].r, arr[0].alpha_dir, arr[0].beta_dir, arr[0].alpha_ex,
                                               arr[0].beta_ex, 0, 0)
    d_sj = struct_array_to_device(sj, dev)
    d_tabs = torch.as_tensor(np.asarray(tp, dtype=np.int64), device=dev)
    d_oidx = torch.as_tensor(np.asarray(order, dtype=np.int32), device=dev)
    out = torch.full((n, 2, L_max + 1), float("nan"), dtype=torch.float64, device=dev)
    check(lib.dwba_il_semisep_lanes_stage_f64(len(order), ptr(d_sj), ptr(d_tabs), ptr(d_oidx), L_max, lbv,
                                              ptr(out), stream_handle(stream)), "dwba_il_semisep_lanes_stage_f64")
    out._dwba_keep = (d_sj, d_tabs, d_oidx, tabs, d_grids)
    return out


def il_plain_f64_batch(jobs, L_max: int, stream=None):
    """Batched plain FP64 contraction.  `jobs` is a list of tuples
    (r, alpha_dir, beta_dir, alpha_ex, beta_ex, n_rows, n_cols) of CUDA tensors.
    Returns a (n_jobs, 2, L_max+1) float64 CUDA tensor."""
    ensure_init()
    lib = load_library()
    n = len(jobs)
    if n == 0:
        raise ValueError("il_plain_f64_batch: no jobs")
    dev = jobs[0][0].device
    arr = (PlainJob * n)()
    nr = (ctypes.c_int32 * n)()
    nc = (ctypes.c_int32 * n)()
    for k, (r, ad, bd, ae, be, n_rows, n_cols) in enumerate(jobs):
        for nm, t, m in (("r", r, n_cols), ("alpha_dir", ad, n_rows), ("beta_dir", bd, n_cols),
                         ("alpha_ex", ae, n_rows), ("beta_ex", be, n_cols)):
            _require_f64_cuda(f"job {k} {nm}", t, m)
        arr[k] = PlainJob(ptr(r), ptr(ad), ptr(bd), ptr(ae), ptr(be), int(n_rows), int(n_cols))
        nr[k] = int(n_rows)
        nc[k] = int(n_cols)
    d_jobs = struct_array_to_device(arr, dev)
    ws_bytes = lib.dwba_il_plain_f64_batch_workspace_bytes(n, nr, nc, L_max)
    if ws_bytes == 0:
        raise ValueError("il_plain_f64_batch: invalid shapes")
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
    out = torch.empty((n, 2, L_max + 1), dtype=torch.float64, device=dev)
    check(lib.dwba_il_plain_f64_batch(n, ptr(d_jobs), nr, nc, L_max, ptr(out), ptr(ws), ws_bytes,
                                      stream_handle(stream)), "dwba_il_plain_f64_batch")
    out._dwba_keep = (ws, d_jobs)
    return out
