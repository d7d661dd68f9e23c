"""Oracle: gradient of the plain contraction with respect to the densities.

Test infrastructure only.  The forward quantity is the plain branch of
radial_ME_all_L (dwba_matrix_elements.py:1243-1245, :1288-1289) for pairs p
sharing one grid:

    I_L(p) = sum_ij w_i rho_a[p, i] K_L(i, j) w_j rho_b[p, j].

It is bilinear, so for the loss sum_{p,L} g[p, L] I_L(p)

    dloss / d rho_a[p, i] = w_i sum_L g[p, L] (K_L (w rho_b[p]))_i
    dloss / d rho_b[p, i] = w_i sum_L g[p, L] (K_L (w rho_a[p]))_i      (K_L symmetric).

K_L x is evaluated exactly in O(N) per L with the semiseparable prefix sums of
semisep.py, in long double.  The reference has no gradient to compare with;
this module is pinned by (tests/test_oracle_pins.py) dense K_L x products on
small grids and central finite differences of semisep.il_all_L, which are exact
for a bilinear form up to rounding.
"""
from __future__ import annotations

import numpy as np

from .semisep import LD, diag_factor


def kmatvec_all_L(r, x, L_max: int) -> np.ndarray:
    """Y[L] = K_L x for L = 0..L_max in long double.  x: (N,) or (N, P);
    result (L_max+1, N) or (L_max+1, N, P), long double."""
    r = np.asarray(r, dtype=np.float64).astype(LD)
    xx = np.asarray(x, dtype=np.float64).astype(LD)
    vec = xx.ndim == 1
    if vec:
        xx = xx[:, None]
    lr = np.log(r)
    out = np.zeros((L_max + 1,) + xx.shape, dtype=LD)
    for L in range(L_max + 1):
        pL = np.exp(L * lr)[:, None]            # r^L
        mL = np.exp(-(L + 1) * lr)[:, None]     # r^-(L+1)
        tp = xx * pL
        P = np.cumsum(tp, axis=0) - tp          # sum_{j<i} x_j r_j^L
        tm = xx * mL
        Q = np.cumsum(tm[::-1], axis=0)[::-1] - tm   # sum_{j>i} x_j r_j^-(L+1)
        out[L] = mL * P + pL * Q + xx * LD(diag_factor(L)) / r[:, None]
    return out[:, :, 0] if vec else out


def density_gradients(r, w, rho_a, rho_b, g, L_max: int, absolute: bool = False):
    """(grad_a, grad_b), each (P, N) float64, of sum_{p,L} g[p,L] I_L(p).
    absolute=True returns the cancellation-free magnitudes
    w_i sum_L |g[p,L]| (K_L |w rho|)_i instead (error scale of a kernel)."""
    r = np.asarray(r, dtype=np.float64)
    w = np.asarray(w, dtype=np.float64)
    ra = np.atleast_2d(np.asarray(rho_a, dtype=np.float64))
    rb = np.atleast_2d(np.asarray(rho_b, dtype=np.float64))
    g = np.atleast_2d(np.asarray(g, dtype=np.float64))
    a = (w[None, :] * ra).T                      # (N, P)
    b = (w[None, :] * rb).T
    if absolute:
        a, b, g = np.abs(a), np.abs(b), np.abs(g)
    Ya = kmatvec_all_L(r, b, L_max)              # (L+1, N, P): K_L beta
    Yb = kmatvec_all_L(r, a, L_max)
    gl = g.T.astype(LD)[:, None, :]              # (L+1, 1, P)
    ga = (w.astype(LD)[:, None] * np.sum(gl * Ya, axis=0)).T
    gb = (w.astype(LD)[:, None] * np.sum(gl * Yb, axis=0)).T
    return ga.astype(np.float64), gb.astype(np.float64)


def plain_forward(r, w, rho_a, rho_b, L_max: int) -> np.ndarray:
    """I[p, L] of the plain contraction (long double evaluation), (P, L_max+1)."""
    from .semisep import il_all_L
    w = np.asarray(w, dtype=np.float64)
    ra = np.atleast_2d(np.asarray(rho_a, dtype=np.float64))
    rb = np.atleast_2d(np.asarray(rho_b, dtype=np.float64))
    return il_all_L(r, (w[None, :] * ra).T, (w[None, :] * rb).T, L_max).T
