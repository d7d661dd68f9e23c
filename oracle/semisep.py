"""Oracle: exact O(N) form of the plain contraction I_L = alpha^T K_L beta.

Test infrastructure only.  K_L(i, j) = r_<^L / r_>^(L+1) off the diagonal and
(1 - 1e-12)^L / r_i on it, which is what the reference builds as
inv_gtr * exp(L * log(min(ratio, 1 - 1e-12) + 1e-300))
(dwba_matrix_elements.py:1057-1080, :1113-1116).  Off the diagonal K_L is
semiseparable: for i < j, K_L(i, j) = r_i^L * r_j^-(L+1), hence

  alpha^T K_L beta = sum_i alpha_i [ r_i^-(L+1) P_i + r_i^L Q_i + beta_i c_L / r_i ]
  P_i = sum_{j<i} beta_j r_j^L,  Q_i = sum_{j>i} beta_j r_j^-(L+1),
  c_L = exp(L * log(1 - 1e-12 + 1e-300)).

Evaluated in numpy long double (64-bit significand) so that it can serve as
the ground truth of FP64 kernels.  Pinned by mpmath brute force (brute.py) and
by hydrogenic closed forms.
"""
from __future__ import annotations

import numpy as np

LD = np.longdouble


def diag_factor(L: int) -> float:
    """(1 - 1e-12)^L as the reference evaluates it on the diagonal."""
    if L == 0:
        return 1.0
    return float(np.exp(L * np.log(min(1.0, 1.0 - 1e-12) + 1e-300)))


def il_all_L(r, alpha, beta, L_max: int, L_min: int = 0) -> np.ndarray:
    """I_L = alpha^T K_L beta for L = L_min..L_max (float64 result of a long
    double evaluation).  alpha and beta may have shape (N,) or (N, P); the
    result has shape (L_max-L_min+1,) or (L_max-L_min+1, P)."""
    r = np.asarray(r, dtype=np.float64).astype(LD)
    a = np.asarray(alpha, dtype=np.float64).astype(LD)
    b = np.asarray(beta, dtype=np.float64).astype(LD)
    vec = a.ndim == 1
    if vec:
        a = a[:, None]
        b = b[:, None]
    out = np.zeros((L_max - L_min + 1, a.shape[1]), dtype=np.float64)
    lr = np.log(r)
    for L in range(L_min, L_max + 1):
        pL = np.exp(L * lr)[:, None]            # r^L
        mL = np.exp(-(L + 1) * lr)[:, None]     # r^-(L+1)
        tb = b * pL
        P = np.cumsum(tb, axis=0) - tb          # sum_{j<i}
        tq = b * mL
        Q = np.cumsum(tq[::-1], axis=0)[::-1] - tq  # sum_{j>i}
        d = LD(diag_factor(L))
        inner = mL * P + pL * Q + b * d / r[:, None]
        out[L - L_min] = np.sum(a * inner, axis=0).astype(np.float64)
    return out[:, 0] if vec else out


def abs_bound_all_L(r, alpha, beta, L_max: int) -> np.ndarray:
    """sum_ij |alpha_i| K_L(i,j) |beta_j|, the cancellation-free magnitude used
    for cancellation-aware relative errors."""
    return il_all_L(r, np.abs(alpha), np.abs(beta), L_max)


def k_dense(r, L: int) -> np.ndarray:
    """Dense K_L for small N (float64)."""
    r = np.asarray(r, dtype=np.float64)
    rl = np.minimum(r[:, None], r[None, :])
    rg = np.maximum(r[:, None], r[None, :])
    ratio = np.minimum(rl / rg, 1.0 - 1e-12)
    inv = 1.0 / rg
    if L == 0:
        return inv
    return inv * np.exp(L * np.log(ratio + 1e-300))
