"""Oracle: radial grid and quadrature weights (test infrastructure only).

Restates grid.py:419-611 of the reference.
"""
from __future__ import annotations

import numpy as np


def trapz_weights(r: np.ndarray) -> np.ndarray:
    """Non-uniform trapezoid weights (grid.py:419-460)."""
    r = np.asarray(r, dtype=np.float64)
    n = r.size
    if n < 2:
        raise ValueError("Need at least 2 grid points for trapezoidal weights.")
    if not np.all(np.diff(r) > 0):
        raise ValueError("r grid must be strictly increasing.")
    w = np.zeros_like(r)
    w[0] = 0.5 * (r[1] - r[0])
    w[-1] = 0.5 * (r[-1] - r[-2])
    if n > 2:
        w[1:-1] = 0.5 * (r[2:] - r[:-2])
    return w


def simpson_weights(r: np.ndarray) -> np.ndarray:
    """Non-uniform Simpson weights over index triplets (grid.py:462-505).

    Triplet (i, i+1, i+2) for i = 0, 2, 4, ... contributes
    c0 = (h1+h2)/6 (2 - h2/h1), c1 = (h1+h2)/6 (h1+h2)^2/(h1 h2),
    c2 = (h1+h2)/6 (2 - h1/h2); for even N the last interval gets a trapezoid.
    The reference accumulates with a Python loop; the sums below add the same
    terms in the same order (0 + c2_prev + c0_next), so the result is
    bit-identical.
    """
    r = np.asarray(r, dtype=np.float64)
    n = r.size
    if n < 3:
        return trapz_weights(r)
    d = np.diff(r)
    i = np.arange(0, n - 2, 2)
    h1 = d[i]
    h2 = d[i + 1]
    s6 = (h1 + h2) / 6.0
    c0 = s6 * (2.0 - h2 / h1)
    c1 = s6 * ((h1 + h2) ** 2 / (h1 * h2))
    c2 = s6 * (2.0 - h1 / h2)
    w = np.zeros(n)
    # Loop order of the reference: w[i] += c0; w[i+1] += c1; w[i+2] += c2,
    # triplet by triplet.  Index i (even, >0) first receives c2 of the
    # previous triplet, then c0 of its own.
    w[i + 2] += c2
    w[i] += c0          # w[i] = c2_prev + c0 (commutative, exact)
    w[i + 1] += c1
    if n % 2 == 0:
        h_last = d[-1]
        w[-2] += 0.5 * h_last
        w[-1] += 0.5 * h_last
    return w


def make_r_grid(r_min: float = 1e-5, r_max: float = 200.0, n_points: int = 4000):
    """Exponential grid r_i = r_min (r_max/r_min)^(i/(N-1)) (grid.py:509-611).

    Returns (r, w_trapz, w_simpson).
    """
    if r_min <= 0.0:
        raise ValueError("r_min must be > 0 to avoid singularities at r=0.")
    if r_max <= r_min:
        raise ValueError("r_max must be > r_min.")
    if n_points < 2:
        raise ValueError("Need at least 2 grid points.")
    idx = np.linspace(0.0, 1.0, n_points)
    r = r_min * ((r_max / r_min) ** idx)
    return r, trapz_weights(r), simpson_weights(r)


def exp_grid(n: int, r_min: float = 1e-5, r_max: float = 1e3) -> np.ndarray:
    """Benchmark grid r_j = r_min exp(j h), h = ln(r_max/r_min)/(N-1)."""
    h = np.log(r_max / r_min) / (n - 1)
    return r_min * np.exp(np.arange(n) * h)
