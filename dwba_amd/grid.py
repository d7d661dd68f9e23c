"""Radial grid and quadrature weights (host side).

Mirrors the reference's ``RadialGrid`` / ``make_r_grid`` (grid.py:40-64,
:509-611) so that callers without the reference can build inputs; any object
with ``r``, ``w_simpson`` and ``w_trapz`` arrays is accepted by the drop-in.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass(frozen=True)
class RadialGrid:
    """r (N,), dr (N,), w_trapz (N,), w_simpson (N,) -- grid.py:40-64."""
    r: np.ndarray
    dr: np.ndarray
    w_trapz: np.ndarray
    w_simpson: np.ndarray


def trapezoid_weights(r: np.ndarray) -> np.ndarray:
    """w_0 = (r_1-r_0)/2, w_i = (r_{i+1}-r_{i-1})/2, w_{N-1} = (r_{N-1}-r_{N-2})/2."""
    r = np.asarray(r, dtype=np.float64)
    if r.size < 2:
        raise ValueError("Need at least 2 grid points for trapezoidal weights.")
    if not np.all(np.diff(r) > 0):
        raise ValueError("r grid must be strictly increasing.")
    w = np.empty_like(r)
    w[0] = 0.5 * (r[1] - r[0])
    w[-1] = 0.5 * (r[-1] - r[-2])
    w[1:-1] = 0.5 * (r[2:] - r[:-2])
    return w


def simpson_weights(r: np.ndarray) -> np.ndarray:
    """Composite non-uniform Simpson weights on consecutive point triplets
    (0,1,2), (2,3,4), ...; with an even point count the last interval is a
    trapezoid.  Per triplet with spacings h1, h2 and s = (h1+h2)/6:
    (s (2 - h2/h1), s (h1+h2)^2/(h1 h2), s (2 - h1/h2)).  The shared end
    point of two triplets receives (0 + c2) + c0, in that order."""
    r = np.asarray(r, dtype=np.float64)
    n = r.size
    if n < 3:
        return trapezoid_weights(r)
    h = np.diff(r)
    h1, h2 = h[0:n - 2:2], h[1:n - 1:2]
    s = (h1 + h2) / 6.0
    first = s * (2.0 - h2 / h1)
    mid = s * ((h1 + h2) ** 2 / (h1 * h2))
    last = s * (2.0 - h1 / h2)
    w = np.zeros(n)
    k = np.arange(0, n - 2, 2)
    w[k + 2] = last
    w[k] += first
    w[k + 1] = mid
    if n % 2 == 0:
        w[-2] += 0.5 * h[-1]
        w[-1] += 0.5 * h[-1]
    return w


def _finish(r: np.ndarray) -> RadialGrid:
    dr = np.empty_like(r)
    dr[:-1] = r[1:] - r[:-1]
    dr[-1] = dr[-2]
    return RadialGrid(r=r, dr=dr, w_trapz=trapezoid_weights(r), w_simpson=simpson_weights(r))


def make_r_grid(r_min: float = 1e-5, r_max: float = 200.0, n_points: int = 4000) -> RadialGrid:
    """Geometric grid r_i = r_min (r_max/r_min)^(i/(N-1))."""
    if r_min <= 0.0:
        raise ValueError("r_min must be > 0 to avoid singularities at r=0.")
    if r_max <= r_min:
        raise ValueError("r_max must be > r_min.")
    if n_points < 2:
        raise ValueError("Need at least 2 grid points.")
    r = r_min * ((r_max / r_min) ** np.linspace(0.0, 1.0, n_points))
    return _finish(r)


def exp_grid(n_points: int, r_min: float = 1e-5, r_max: float = 1e3) -> RadialGrid:
    """Benchmark grid r_j = r_min exp(j h), h = ln(r_max/r_min)/(N-1)."""
    h = np.log(r_max / r_min) / (n_points - 1)
    return _finish(r_min * np.exp(np.arange(n_points) * h))
