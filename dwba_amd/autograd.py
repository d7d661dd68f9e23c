"""Differentiable plain multipole integrals (potential fitting).

``multipole_integrals(rho_a, rho_b, r, w, L_max, precision)`` returns, for P
channel pairs on one radial grid,

    I[p, L] = sum_ij w_i rho_a[p, i] K_L(i, j) w_j rho_b[p, j],   L = 0..L_max,

the plain branch of ``radial_ME_all_L`` (dwba_matrix_elements.py:1243-1245,
:1288-1289) with K_L = r_<^L / r_>^(L+1) and the reference's diagonal clamp.
It is a ``torch.autograd.Function``; nothing of size N x N is ever stored.
The reference has no gradient; this is new functionality for fitting
potentials through the integrals.

precision / contraction:
  "fp64", "semiseparable" (default)  forward through the O(N) semiseparable
          kernel (dwba_il_semisep_f64), backward through
          dwba_il_backward_semisep_f64: K_L (w rho) by the same O(N)
          recurrences for every L, grad = w sum_L g[p, L] K_L (w rho), FP64;
  "fp64", "dense"  forward through the dense FP64 plain kernel, backward
          through dwba_il_backward_f64, which regenerates
          Keff_p = sum_L g[p, L] K_L element by element (Horner in q = r_< / r_>);
  "fp32"  forward through the stacked FP32 MFMA GEMM (FP32 operands and
          accumulation, FP64 alpha epilogue), backward with Keff and the
          products in FP32, 256-column chunks summed in FP64 (dense).
Gradients flow to rho_a and rho_b only (r and w are constants).
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _hip

XROWS, XCOLS, XK = 256, 128, 32


def _check_inputs(rho_a, rho_b, r, w, L_max, precision, contraction="semiseparable"):
    if precision not in ("fp64", "fp32"):
        raise ValueError(f"precision must be 'fp64' or 'fp32', got {precision!r}")
    if contraction not in ("semiseparable", "dense"):
        raise ValueError(f"contraction must be 'semiseparable' or 'dense', got {contraction!r}")
    if not 0 <= int(L_max) <= 63:
        raise ValueError("L_max must be in [0, 63]")
    for nm, t in (("rho_a", rho_a), ("rho_b", rho_b), ("r", r), ("w", w)):
        if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float64):
            raise ValueError(f"{nm} must be a float64 CUDA tensor")
    if rho_a.dim() != 2 or rho_a.shape != rho_b.shape:
        raise ValueError("rho_a and rho_b must both have shape (P, N)")
    n = rho_a.shape[1]
    if r.shape != (n,) or w.shape != (n,):
        raise ValueError(f"r and w must have shape ({n},)")
    if n < 2 or rho_a.shape[0] < 1:
        raise ValueError("need N >= 2 grid points and P >= 1 pairs")


def _weighted(lib, w, rho, ld, sh):
    P, n = rho.shape
    out = torch.empty((P, ld), dtype=torch.float64, device=rho.device)
    _hip.check(lib.dwba_weight_rows(_hip.ptr(w), n, P, _hip.ptr(rho), n, _hip.ptr(out), ld, sh), "dwba_weight_rows")
    return out


def _forward_fp64(lib, rho_a, rho_b, r, w, L_max, sh):
    P, n = rho_a.shape
    A = _weighted(lib, w, rho_a, n, sh)
    B = _weighted(lib, w, rho_b, n, sh)
    zero = torch.zeros(n, dtype=torch.float64, device=rho_a.device)
    jobs = []
    for q in range(0, P, 2):
        a1, b1 = (A[q + 1], B[q + 1]) if q + 1 < P else (zero, zero)
        jobs.append((r, A[q], B[q], a1, b1, n, n))
    out = _hip.il_plain_f64_batch(jobs, L_max)          # (jobs, 2, L+1): slot 0 = pair 2q, slot 1 = 2q+1
    return out.reshape(-1, L_max + 1)[:P].clone()


def _forward_semisep(lib, rho_a, rho_b, r, w, L_max, sh):
    P, n = rho_a.shape
    A = _weighted(lib, w, rho_a, n, sh)
    B = _weighted(lib, w, rho_b, n, sh)
    zero = torch.zeros(n, dtype=torch.float64, device=rho_a.device)
    jobs = []
    for q in range(0, P, 2):
        a1, b1 = (A[q + 1], B[q + 1]) if q + 1 < P else (zero, zero)
        jobs.append((r, A[q], B[q], a1, b1, n, n))
    out = _hip.il_semisep_f64_batch(jobs, L_max)        # (jobs, 2, L+1): slot 0 = pair 2q, slot 1 = 2q+1
    return out.reshape(-1, L_max + 1)[:P].clone()


def _backward(lib, r, w, rho_a, rho_b, g, L_max, precision, contraction, ga, gb, sh):
    P, n = rho_a.shape
    if precision == "fp64" and contraction == "semiseparable":
        nb = lib.dwba_il_backward_semisep_f64_workspace_bytes(n, P, L_max)
        ws = torch.empty(max(int(nb), 1), dtype=torch.uint8, device=rho_a.device)
        _hip.check(lib.dwba_il_backward_semisep_f64(_hip.ptr(r), _hip.ptr(w), n, P, n, _hip.ptr(rho_a),
                                                    _hip.ptr(rho_b), _hip.ptr(g), L_max, _hip.ptr(ga), _hip.ptr(gb),
                                                    _hip.ptr(ws), nb, sh), "dwba_il_backward_semisep_f64")
        return ws
    fn = lib.dwba_il_backward_f64 if precision == "fp64" else lib.dwba_il_backward_f32
    _hip.check(fn(_hip.ptr(r), _hip.ptr(w), n, P, n, _hip.ptr(rho_a), _hip.ptr(rho_b), _hip.ptr(g), L_max,
                  _hip.ptr(ga), _hip.ptr(gb), sh), "dwba_il_backward")
    return None


def _forward_fp32(lib, rho_a, rho_b, r, w, L_max, sh):
    from .batch import XgemmEnergy
    P, n = rho_a.shape
    dev = rho_a.device
    ld = (n + XK - 1) // XK * XK
    r_h = r.detach().cpu().numpy()
    logr = np.log(np.maximum(r_h, 1e-300))
    n_rb = (n + XROWS - 1) // XROWS
    for rb in range(n_rb):
        a, b = rb * XROWS, min(n, rb * XROWS + XROWS) - 1
        if L_max * (logr[b] - logr[a]) > 80.0:
            raise ValueError("grid too coarse for the FP32 factorised kernel; use precision='fp64'")
    logr_d = torch.as_tensor(logr, device=dev)
    A = _weighted(lib, w, rho_a, ld, sh)
    B = _weighted(lib, w, rho_b, ld, sh)
    Bf = torch.empty((P, ld), dtype=torch.float32, device=dev)
    _hip.check(lib.dwba_cast_f64_f32(_hip.ptr(B), _hip.ptr(Bf), B.numel(), sh), "dwba_cast_f64_f32")
    n_cb = (P + XCOLS - 1) // XCOLS
    ce = torch.full((n_cb,), n, dtype=torch.int32, device=dev)
    rbk = torch.full((n_cb,), n_rb, dtype=torch.int32, device=dev)
    desc = (XgemmEnergy * 1)(XgemmEnergy(_hip.ptr(logr_d), _hip.ptr(Bf), 0, 0, _hip.ptr(A), _hip.ptr(ce),
                                         _hip.ptr(rbk), 0, n, P, ld, n_cb, n_rb, 0, 0, 0))
    desc_d = _hip.struct_array_to_device(desc, dev)
    items = np.array([(0, L, cb, rb) for rb in range(n_rb) for cb in range(n_cb) for L in range(L_max + 1)],
                     dtype=np.int32)
    items_d = torch.as_tensor(items, device=dev)
    partial = torch.zeros(int(lib.dwba_xgemm_partial_doubles(n_rb, n_cb, L_max)), dtype=torch.float64, device=dev)
    counter = torch.zeros(1, dtype=torch.int32, device=dev)
    out = torch.zeros((P, 2, L_max + 1), dtype=torch.float64, device=dev)
    _hip.check(lib.dwba_xgemm_f32(1, _hip.ptr(desc_d), len(items), _hip.ptr(items_d), L_max, _hip.ptr(partial),
                                  _hip.ptr(counter), n_cb, _hip.ptr(out), 0, sh), "dwba_xgemm_f32")
    res = out[:, 0, :].clone()
    res._dwba_keep = (A, Bf, ce, rbk, desc_d, items_d, partial, counter, logr_d)
    return res


class MultipoleIntegrals(torch.autograd.Function):
    """I[p, L] = (w rho_a[p])^T K_L (w rho_b[p]); backward via the HIP kernel."""

    @staticmethod
    def forward(ctx, rho_a, rho_b, r, w, L_max: int, precision: str = "fp64", contraction: str = "semiseparable"):
        _check_inputs(rho_a, rho_b, r, w, L_max, precision, contraction)
        _hip.ensure_init(rho_a.device.index)
        lib = _hip.load_library()
        sh = _hip.stream_handle()
        rho_a = rho_a.detach().contiguous()
        rho_b = rho_b.detach().contiguous()
        r = r.detach().contiguous()
        w = w.detach().contiguous()
        if precision == "fp32":
            fwd = _forward_fp32
        else:
            fwd = _forward_semisep if contraction == "semiseparable" else _forward_fp64
        out = fwd(lib, rho_a, rho_b, r, w, int(L_max), sh)
        ctx.save_for_backward(rho_a, rho_b, r, w)
        ctx.L_max = int(L_max)
        ctx.precision = precision
        ctx.contraction = contraction
        return out

    @staticmethod
    def backward(ctx, grad_out):
        rho_a, rho_b, r, w = ctx.saved_tensors
        need_a, need_b = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
        if not (need_a or need_b):
            return None, None, None, None, None, None
        lib = _hip.load_library()
        P, n = rho_a.shape
        g = grad_out.detach().to(torch.float64).contiguous()
        ga = torch.empty((P, n), dtype=torch.float64, device=rho_a.device) if need_a else None
        gb = torch.empty((P, n), dtype=torch.float64, device=rho_a.device) if need_b else None
        ws = _backward(lib, r, w, rho_a, rho_b, g, ctx.L_max, ctx.precision, ctx.contraction, ga, gb,
                       _hip.stream_handle())
        if ws is not None:
            (ga if ga is not None else gb)._dwba_ws = ws     # the workspace lives until the stream is done with it
        return ga, gb, None, None, None, None, None


def multipole_integrals(rho_a, rho_b, r, w, L_max: int, precision: str = "fp64", contraction: str = "semiseparable"):
    """Differentiable I[p, L] for P pairs on one grid; see the module docstring."""
    return MultipoleIntegrals.apply(rho_a, rho_b, r, w, L_max, precision, contraction)


def density_gradients(rho_a, rho_b, r, w, g, L_max: int, precision: str = "fp64", contraction: str = "semiseparable"):
    """Gradients of sum_{p,L} g[p, L] I[p, L] with respect to rho_a and rho_b
    directly (no autograd graph).  Returns (grad_a, grad_b), each (P, N)."""
    _check_inputs(rho_a, rho_b, r, w, L_max, precision, contraction)
    _hip.ensure_init(rho_a.device.index)
    lib = _hip.load_library()
    rho_a, rho_b = rho_a.contiguous(), rho_b.contiguous()
    P, n = rho_a.shape
    g = g.to(torch.float64).contiguous()
    if g.shape != (P, L_max + 1):
        raise ValueError(f"g must have shape ({P}, {L_max + 1})")
    ga = torch.empty((P, n), dtype=torch.float64, device=rho_a.device)
    gb = torch.empty_like(ga)
    ws = _backward(lib, r.contiguous(), w.contiguous(), rho_a, rho_b, g, L_max, precision, contraction, ga, gb,
                   _hip.stream_handle())
    if ws is not None:
        ga._dwba_ws = ws
    return ga, gb
