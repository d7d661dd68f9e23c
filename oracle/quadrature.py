"""Oracle: Clenshaw-Curtis weights, phase nodes, log-r interpolation.

Test infrastructure only.  Restates oscillatory_integrals.py:120-171 (CC
reference weights), :237-302 (log-r interpolation), :1456-1509 (phase
sampling check), :1539-1618 (clenshaw_curtis_nodes), :1625-1713 (phase
nodes) and :1716-1755 (adaptive phase increment).
"""
from __future__ import annotations

import math

import numpy as np


def cc_module_weights(n: int):
    """Module-level 'CC' weights on [-1, 1] (oscillatory_integrals.py:120-142).

    Built for n = 5 only in the reference.  The trailing j = j_max term is
    added twice (once in the sum, once more with factor 1/2), so the weights
    sum to 2.0667 for n = 5 instead of 2: kept as is.
    """
    theta = np.pi * np.arange(n) / (n - 1)
    x = np.cos(theta)
    j_max = (n - 1) // 2
    j_vals = np.arange(1, j_max + 1)
    if j_max > 0:
        cos_terms = np.cos(2 * j_vals[:, None] * theta[None, :])
        denom = 4 * j_vals ** 2 - 1
        ws = np.sum(cos_terms / denom[:, None], axis=0)
        if (n - 1) % 2 == 0:
            ws += 0.5 * np.cos(2 * j_max * theta) / (4 * j_max ** 2 - 1)
    else:
        ws = np.zeros(n)
    w = (2.0 / (n - 1)) * (1 - 2 * ws)
    return x, w


def clenshaw_curtis_nodes(n: int, a: float, b: float):
    """clenshaw_curtis_nodes (oscillatory_integrals.py:1539-1618).

    Nodes run from b down to a.  Weights sum to (b-a) and integrate
    polynomials of degree <= n-2 exactly.
    """
    if n < 2:
        raise ValueError("Clenshaw-Curtis requires n >= 2 nodes")
    if n == 2:
        return np.array([b, a]), np.array([0.5, 0.5]) * (b - a)
    theta = np.pi * np.arange(n) / (n - 1)
    x = np.cos(theta)
    nodes = 0.5 * (b - a) * (x + 1) + a
    N = n - 1
    j_max = N // 2
    w = np.ones(n)
    if j_max > 0:
        j_vals = np.arange(1, j_max + 1)
        b_j = 2.0 / (1 - 4 * j_vals ** 2)
        w += np.sum(b_j[:, None] * np.cos(2 * j_vals[:, None] * theta[None, :]), axis=0)
    if N % 2 == 0:
        w += np.cos(N * theta) / (1 - N ** 2)
    w *= 2.0 / N
    w[0] /= 2
    w[-1] /= 2
    w *= 0.5 * (b - a)
    return nodes, w


def cc_ref(n: int):
    """_get_cc_ref (oscillatory_integrals.py:159-171): n == 5 uses the module
    formula, every other n uses clenshaw_curtis_nodes(n, -1, 1)."""
    if n < 2:
        raise ValueError("Clenshaw-Curtis requires n_nodes >= 2.")
    if n == 5:
        return cc_module_weights(5)
    return clenshaw_curtis_nodes(n, -1.0, 1.0)


def check_phase_sampling(r, k_total, threshold=np.pi / 4, eta_total=0.0):
    """check_phase_sampling (oscillatory_integrals.py:1456-1509)."""
    if k_total < 1e-10:
        return 0.0, True, -1
    r = np.asarray(r, dtype=np.float64)
    dr = np.diff(r)
    if abs(eta_total) > 1e-10:
        r_mid = np.maximum(0.5 * (r[:-1] + r[1:]), 1e-10)
        pps = np.abs((k_total + eta_total / r_mid) * dr)
    else:
        pps = np.abs(k_total * dr)
    max_phase = float(np.max(pps))
    mask = pps > threshold
    if np.any(mask):
        return max_phase, False, int(np.argmax(mask))
    return max_phase, True, -1


def adaptive_phase_increment(r_grid, k_total, requested, eta_total=0.0,
                             threshold=np.pi / 4, min_increment=np.pi / 16):
    """_adaptive_phase_increment_for_grid (oscillatory_integrals.py:1716-1755)."""
    dphi_req = float(max(min_increment, requested))
    dphi_req = min(dphi_req, threshold)
    if k_total < 1e-10 or len(r_grid) < 2:
        return dphi_req
    max_phase, ok, _ = check_phase_sampling(np.asarray(r_grid), float(k_total),
                                            threshold=float(threshold),
                                            eta_total=float(eta_total))
    if ok:
        return dphi_req
    scale = float(threshold) / max(float(max_phase), 1e-12)
    dphi_eff = dphi_req * max(0.25, min(1.0, scale))
    return float(np.clip(dphi_eff, min_increment, threshold))


def generate_phase_nodes(r_start, r_end, k_total, phase_increment=np.pi / 2, eta_total=0.0):
    """generate_phase_nodes (oscillatory_integrals.py:1625-1713)."""
    if k_total < 1e-10:
        return np.array([r_start, r_end])
    if abs(eta_total) < 1e-10:
        dr = phase_increment / k_total
        n_points = int(np.ceil((r_end - r_start) / dr)) + 1
        return np.linspace(r_start, r_end, n_points)
    nodes = [r_start]
    r = r_start
    while r < r_end:
        phi_prime = k_total + eta_total / max(r, 1e-10)
        dr = phase_increment / abs(phi_prime)
        dr = min(dr, (r_end - r_start) / 10)
        dr = max(dr, 1e-6)
        r_next = r + dr
        if r_next >= r_end:
            r_next = r_end
        nodes.append(r_next)
        r = r_next
        if len(nodes) > 10000:
            break
    return np.array(nodes)


def is_log_spaced_grid(r, rel_tol=5e-3) -> bool:
    """_is_log_spaced_grid (oscillatory_integrals.py:237-255)."""
    r = np.asarray(r, dtype=np.float64)
    if r.ndim != 1 or len(r) < 3:
        return False
    if r[0] <= 0.0 or np.any(np.diff(r) <= 0.0):
        return False
    dlog = np.diff(np.log(r))
    mean = float(np.mean(dlog))
    if not np.isfinite(mean) or mean <= 0.0:
        return False
    return float(np.max(np.abs(dlog - mean))) <= rel_tol * mean


def interp_on_radial_grid(x_new, x, y):
    """_interp_on_radial_grid (oscillatory_integrals.py:279-302): np.interp in
    log r on log-spaced grids, in r otherwise; x_new clipped to [x0, x_end]."""
    x = np.asarray(x, dtype=np.float64)
    x_new = np.asarray(x_new, dtype=np.float64)
    y = np.asarray(y)
    xc = np.clip(x_new, x[0], x[-1])
    if np.iscomplexobj(y):
        return interp_on_radial_grid(xc, x, y.real) + 1j * interp_on_radial_grid(xc, x, y.imag)
    if is_log_spaced_grid(x):
        xs = np.maximum(x, 1e-300)
        xns = np.maximum(xc, xs[0])
        return np.interp(np.log(xns), np.log(xs), y)
    return np.interp(xc, x, y)


def fsum(values) -> float:
    """math.fsum, the correctly rounded sum used by _kahan_sum_real/complex
    (oscillatory_integrals.py:178-218)."""
    return math.fsum(values)
