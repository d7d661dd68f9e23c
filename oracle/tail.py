"""Oracle: outer oscillatory tail I_out of the direct integral.

Test infrastructure only.  Restates, step by step, the chain that
``radial_ME_all_L`` calls for L >= 1 excitation tails
(dwba_matrix_elements.py:1165-1192):

  dwba_outer_integral_1d            oscillatory_integrals.py:459-535
  _get_dwba_outer_phase_functions   oscillatory_integrals.py:350-456
  compute_outer_integral_oscillatory oscillatory_integrals.py:1361-1451
  filon_oscillatory_integral        oscillatory_integrals.py:1280-1354
  _filon_segment_complex            oscillatory_integrals.py:1165-1277
  levin_oscillatory_integral        oscillatory_integrals.py:1099-1158
  _levin_segment_complex            oscillatory_integrals.py:1024-1096
  _chebyshev_differentiation_matrix oscillatory_integrals.py:970-1017
  _analytical_multipole_tail        oscillatory_integrals.py:2064-2184

Quirks kept on purpose because the reference computes them:
  * Filon uses the constant frequency |Phi'(mid)| (absolute value) and the
    phase offset Phi(r_m) - |Phi'(mid)| r_m;
  * the closed-form Filon beta (theta >= 0.3) is 2(theta(1+cos 2theta) -
    sin 2theta)/theta^3, which is not the textbook coefficient; the Taylor
    branch (theta < 0.3) is textbook;
  * the alpha term carries an extra factor -i/omega;
  * Levin pairs the Chebyshev matrix of descending Lobatto points with
    ascending nodes (a sign flip of d/dr).
"""
from __future__ import annotations

import math
import os

import numpy as np
from scipy.integrate import quad, simpson


def _env_int(name, default, min_value=1):
    raw = os.environ.get(name, "").strip()
    try:
        return max(min_value, int(raw)) if raw else max(min_value, int(default))
    except Exception:
        return max(min_value, int(default))


def _env_float(name, default, min_value=0.0):
    raw = os.environ.get(name, "").strip()
    try:
        return max(min_value, float(raw)) if raw else max(min_value, float(default))
    except Exception:
        return max(min_value, float(default))


def filon_max_segments() -> int:
    """DWBA_MAX_FILON_SEGMENTS (oscillatory_integrals.py:75)."""
    return _env_int("DWBA_MAX_FILON_SEGMENTS", 4096, min_value=128)


def filon_max_effective_dphi() -> float:
    """DWBA_MAX_EFFECTIVE_DPHI (oscillatory_integrals.py:76-80)."""
    return _env_float("DWBA_MAX_EFFECTIVE_DPHI", np.pi / 2, min_value=np.pi / 16)


def _ksum(values) -> complex:
    v = np.asarray(values, dtype=complex)
    return complex(math.fsum(v.real), math.fsum(v.imag))


class PhaseFunctions:
    """The six phase callables of _get_dwba_outer_phase_functions
    (oscillatory_integrals.py:350-456) for the sinA*sinB split
    (compute_product_phases, :312-347)."""

    def __init__(self, k_a, l_a, delta_a, eta_a, sigma_a, k_b, l_b, delta_b, eta_b, sigma_b):
        self.k_a, self.l_a, self.eta_a = k_a, l_a, eta_a
        self.k_b, self.l_b, self.eta_b = k_b, l_b, eta_b
        self.k_minus = k_a - k_b
        self.k_plus = k_a + k_b
        pa = -l_a * np.pi / 2 + sigma_a + delta_a
        pb = -l_b * np.pi / 2 + sigma_b + delta_b
        self.c_minus = pa - pb
        self.c_plus = pa + pb

    def phi(self, sign, r):
        ka, kb, la, lb, ea, eb = self.k_a, self.k_b, self.l_a, self.l_b, self.eta_a, self.eta_b
        if sign < 0:
            res = self.k_minus * r + self.c_minus
            if abs(ea) > 1e-15:
                res = res + ea * np.log(2 * ka * r + 1e-30)
            if abs(eb) > 1e-15:
                res = res - eb * np.log(2 * kb * r + 1e-30)
            if la > 0:
                res = res - (la * (la + 1)) / (2.0 * ka * r)
            if lb > 0:
                res = res + (lb * (lb + 1)) / (2.0 * kb * r)
            return res
        res = self.k_plus * r + self.c_plus
        if abs(ea) > 1e-15:
            res = res + ea * np.log(2 * ka * r + 1e-30)
        if abs(eb) > 1e-15:
            res = res + eb * np.log(2 * kb * r + 1e-30)
        if la > 0:
            res = res - (la * (la + 1)) / (2.0 * ka * r)
        if lb > 0:
            res = res - (lb * (lb + 1)) / (2.0 * kb * r)
        return res

    def dphi(self, sign, r):
        ka, kb, la, lb, ea, eb = self.k_a, self.k_b, self.l_a, self.l_b, self.eta_a, self.eta_b
        s = -1.0 if sign < 0 else 1.0
        res = self.k_minus if sign < 0 else self.k_plus
        if abs(ea) > 1e-15:
            res = res + ea / r
        if abs(eb) > 1e-15:
            res = res + s * eb / r
        if la > 0:
            res = res + (la * (la + 1)) / (2.0 * ka * r ** 2)
        if lb > 0:
            res = res + s * (lb * (lb + 1)) / (2.0 * kb * r ** 2)
        return res

    def d2phi(self, sign, r):
        ka, kb, la, lb, ea, eb = self.k_a, self.k_b, self.l_a, self.l_b, self.eta_a, self.eta_b
        s = -1.0 if sign < 0 else 1.0
        res = 0.0
        if abs(ea) > 1e-15:
            res = res - ea / (r * r)
        if abs(eb) > 1e-15:
            res = res - s * eb / (r * r)
        if la > 0:
            res = res - (la * (la + 1)) / (ka * r ** 3)
        if lb > 0:
            res = res - s * (lb * (lb + 1)) / (kb * r ** 3)
        return res


def filon_segment(f_vals, r_nodes, omega, phase_offset=0.0) -> complex:
    """_filon_segment_complex (oscillatory_integrals.py:1165-1277)."""
    f_vals = np.asarray(f_vals, dtype=float)
    r_nodes = np.asarray(r_nodes, dtype=float)
    n = len(r_nodes)
    if n < 2:
        return 0j
    a, b = r_nodes[0], r_nodes[-1]
    h = b - a
    if h < 1e-15:
        return 0j
    if abs(omega * h) < 1e-3:
        mid_phase = omega * (a + b) / 2 + phase_offset
        return simpson(f_vals, x=r_nodes) * np.exp(1j * mid_phase)
    if n == 3:
        f0, f1, f2 = f_vals
        r0, r1, r2 = r_nodes
        theta = omega * h / 2
        sin_t, cos_t = np.sin(theta), np.cos(theta)
        if abs(theta) < 0.3:
            t2 = theta ** 2
            t3 = theta ** 3
            alpha = 2 * t3 / 45 * (1 - t2 / 7 + t2 ** 2 / 189)
            beta = 2 / 3 + 2 * t2 / 15 - 4 * t2 ** 2 / 105 + 2 * t2 ** 3 / 567
            gamma = 4 / 3 - 2 * t2 / 15 + 1 * t2 ** 2 / 210 - t2 ** 3 / 11340
        else:
            t2 = theta ** 2
            sin_2t, cos_2t = np.sin(2 * theta), np.cos(2 * theta)
            alpha = (t2 + theta * sin_2t / 2 - 2 * sin_t ** 2) / (theta ** 3)
            beta = 2 * (theta * (1 + cos_2t) - sin_2t) / (theta ** 3)
            gamma = 4 * (sin_t - theta * cos_t) / (theta ** 3)
        e0 = np.exp(1j * (omega * r0 + phase_offset))
        e1 = np.exp(1j * (omega * r1 + phase_offset))
        e2 = np.exp(1j * (omega * r2 + phase_offset))
        return (h / 2) * (alpha * (-1j / omega) * (f2 * e2 - f0 * e0)
                          + beta * (f0 * e0 + f2 * e2) + gamma * f1 * e1)
    res = 0j
    for i in range(0, n - 2, 2):
        res += filon_segment(f_vals[i:i + 3], r_nodes[i:i + 3], omega, phase_offset)
    if n > 2 and (n - 1) % 2 == 1:
        rl = r_nodes[-2:]
        fl = f_vals[-2:]
        pl = omega * rl + phase_offset
        hl = rl[-1] - rl[0]
        res += 0.5 * hl * (fl[0] * np.exp(1j * pl[0]) + fl[1] * np.exp(1j * pl[1]))
    return res


def filon_integral(f_func, omega, phase_offset, r_start, r_end, delta_phi=np.pi / 4,
                   n_nodes_per_segment=5) -> complex:
    """filon_oscillatory_integral (oscillatory_integrals.py:1280-1354)."""
    if abs(omega) < 1e-10:
        re, _ = quad(lambda r: f_func(r) * np.cos(phase_offset), r_start, r_end)
        im, _ = quad(lambda r: f_func(r) * np.sin(phase_offset), r_start, r_end)
        return re + 1j * im
    max_seg = filon_max_segments()
    dr_seg = delta_phi / abs(omega)
    n_raw = max(1, int(np.ceil((r_end - r_start) / dr_seg)))
    n_seg = min(n_raw, max_seg)
    eff = abs(omega) * (r_end - r_start) / max(1, n_seg)
    n_eff = int(n_nodes_per_segment)
    dmax = filon_max_effective_dphi()
    if eff > dmax:
        scale = int(np.ceil(eff / dmax))
        n_eff = min(15, max(n_nodes_per_segment, n_nodes_per_segment + 2 * (scale - 1)))
    bounds = np.linspace(r_start, r_end, n_seg + 1)
    f_vec = getattr(f_func, "vec", None)
    if f_vec is not None:
        return _ksum(_filon_segments_vec(f_vec, bounds, n_eff, omega, phase_offset))
    contrib = []
    for i in range(n_seg):
        nodes = np.linspace(bounds[i], bounds[i + 1], n_eff)
        fv = np.array([f_func(x) for x in nodes], dtype=float)
        contrib.append(filon_segment(fv, nodes, omega, phase_offset))
    return _ksum(contrib)


def _filon_segments_vec(f_vec, bounds, n_eff, omega, phase_offset):
    """filon_segment over every segment [bounds[i], bounds[i+1]] at once:
    the same nodes (np.linspace arithmetic: start + i*step, last node = stop),
    the same per-triple Filon coefficients and the same left-to-right sum of
    the triples of a segment; a segment with |omega h| < 1e-3 (Simpson
    branch) or an even node count falls back to filon_segment.  Returns the
    per-segment complex contributions in segment order."""
    a, b = bounds[:-1], bounds[1:]
    step = (b - a) / (n_eff - 1)
    nodes = np.arange(n_eff)[None, :] * step[:, None] + a[:, None]
    nodes[:, -1] = b
    fv = f_vec(nodes)
    h_seg = nodes[:, -1] - nodes[:, 0]
    out = np.zeros(len(a), dtype=complex)
    slow = (h_seg < 1e-15) | (np.abs(omega * h_seg) < 1e-3)
    if n_eff % 2 == 0 or n_eff < 3:
        slow[:] = True
    fast = ~slow
    if np.any(fast):
        R, F = nodes[fast], fv[fast]
        res = np.zeros(R.shape[0], dtype=complex)
        for i in range(0, n_eff - 2, 2):
            r0, r1, r2 = R[:, i], R[:, i + 1], R[:, i + 2]
            f0, f1, f2 = F[:, i], F[:, i + 1], F[:, i + 2]
            h = r2 - r0
            theta = omega * h / 2
            t2 = theta ** 2
            small = np.abs(theta) < 0.3
            with np.errstate(divide="ignore", invalid="ignore"):
                sin_t, cos_t = np.sin(theta), np.cos(theta)
                sin_2t, cos_2t = np.sin(2 * theta), np.cos(2 * theta)
                al_c = (t2 + theta * sin_2t / 2 - 2 * sin_t ** 2) / (theta ** 3)
                be_c = 2 * (theta * (1 + cos_2t) - sin_2t) / (theta ** 3)
                ga_c = 4 * (sin_t - theta * cos_t) / (theta ** 3)
            t3 = theta ** 3
            al_s = 2 * t3 / 45 * (1 - t2 / 7 + t2 ** 2 / 189)
            be_s = 2 / 3 + 2 * t2 / 15 - 4 * t2 ** 2 / 105 + 2 * t2 ** 3 / 567
            ga_s = 4 / 3 - 2 * t2 / 15 + 1 * t2 ** 2 / 210 - t2 ** 3 / 11340
            alpha = np.where(small, al_s, al_c)
            beta = np.where(small, be_s, be_c)
            gamma = np.where(small, ga_s, ga_c)
            e0 = np.exp(1j * (omega * r0 + phase_offset))
            e1 = np.exp(1j * (omega * r1 + phase_offset))
            e2 = np.exp(1j * (omega * r2 + phase_offset))
            tri = (h / 2) * (alpha * (-1j / omega) * (f2 * e2 - f0 * e0) + beta * (f0 * e0 + f2 * e2) + gamma * f1 * e1)
            # a triple of zero width contributes 0 (filon_segment's h < 1e-15 test)
            tri = np.where(h < 1e-15, 0j, tri)
            res = res + tri
        out[fast] = res
    for i in np.nonzero(slow)[0]:
        out[i] = filon_segment(fv[i], nodes[i], omega, phase_offset)
    return out


def chebyshev_diff_matrix(n: int) -> np.ndarray:
    """_chebyshev_differentiation_matrix (oscillatory_integrals.py:970-1017):
    Lobatto points x_i = cos(pi i/(n-1)) in descending order."""
    if n < 2:
        return np.zeros((n, n))
    x = np.cos(np.pi * np.arange(n) / (n - 1))
    c = np.ones(n)
    c[0] = 2.0
    c[-1] = 2.0
    i = np.arange(n)[:, None]
    j = np.arange(n)[None, :]
    sign = (-1.0) ** (i + j)
    xd = x[:, None] - x[None, :]
    np.fill_diagonal(xd, 1.0)
    D = (c[:, None] / c[None, :]) * sign / xd
    np.fill_diagonal(D, 0.0)
    D[np.diag_indices(n)] = -np.sum(D, axis=1)
    return D


def levin_segment(f_vals, r_nodes, phi_vals, dphi_vals) -> complex:
    """_levin_segment_complex (oscillatory_integrals.py:1024-1096)."""
    n = len(r_nodes)
    if n < 2:
        return 0j
    a, b = r_nodes[0], r_nodes[-1]
    h = b - a
    if h < 1e-15:
        return 0j
    D = chebyshev_diff_matrix(n) * (2.0 / h)
    A = D + 1j * np.diag(dphi_vals)
    try:
        u = np.linalg.solve(A, np.asarray(f_vals).astype(complex))
    except np.linalg.LinAlgError:
        u = np.linalg.lstsq(A, np.asarray(f_vals).astype(complex), rcond=None)[0]
    return u[-1] * np.exp(1j * phi_vals[-1]) - u[0] * np.exp(1j * phi_vals[0])


def levin_integral(f_func, phi_func, dphi_func, r_start, r_end, n_nodes=8, n_segments=1) -> complex:
    """levin_oscillatory_integral (oscillatory_integrals.py:1099-1158)."""
    n_segments = max(1, n_segments)
    bounds = np.linspace(r_start, r_end, n_segments + 1)
    x_ref = np.cos(np.pi * np.arange(n_nodes) / (n_nodes - 1))[::-1]
    contrib = []
    for i in range(n_segments):
        a, b = bounds[i], bounds[i + 1]
        nodes = 0.5 * (b - a) * (x_ref + 1) + a
        fv = np.array([f_func(x) for x in nodes], dtype=float)
        contrib.append(levin_segment(fv, nodes, phi_func(nodes), dphi_func(nodes)))
    return _ksum(contrib)


def outer_integral_oscillatory(f_func, phi_func, dphi_func, d2phi_func, r_m, r_max,
                               delta_phi=np.pi / 4, n_nodes=8, filon_threshold=0.1) -> complex:
    """compute_outer_integral_oscillatory (oscillatory_integrals.py:1361-1451)."""
    if r_max <= r_m + 1e-10:
        return 0j
    mid = (r_m + r_max) / 2
    omega_mid = abs(dphi_func(mid))
    if omega_mid < 1e-10:
        re, _ = quad(lambda r: f_func(r) * np.cos(phi_func(r)), r_m, r_max)
        im, _ = quad(lambda r: f_func(r) * np.sin(phi_func(r)), r_m, r_max)
        return re + 1j * im
    h = delta_phi / omega_mid
    if abs(d2phi_func(mid)) * h * h < filon_threshold:
        return filon_integral(f_func, omega_mid, phi_func(r_m) - omega_mid * r_m,
                              r_m, r_max, delta_phi)
    n_seg = max(1, int(np.ceil((r_max - r_m) / h)))
    return levin_integral(f_func, phi_func, dphi_func, r_m, r_max, n_nodes, n_seg)


def multipole_envelope(moment_L: float, L: int):
    """Envelope moment_L / r^(L+1) of radial_ME_all_L (dwba_matrix_elements.py:1170-1175).
    env.vec evaluates it on an array (used by the vectorised Filon sweep)."""
    def env(r):
        return moment_L / (r ** (L + 1)) if r > 1e-6 else 0.0

    def vec(r):
        r = np.asarray(r, dtype=float)
        with np.errstate(divide="ignore", over="ignore"):
            return np.where(r > 1e-6, moment_L / (r ** (L + 1)), 0.0)
    env.vec = vec
    return env


def dwba_outer_integral(envelope, k_a, l_a, delta_a, eta_a, sigma_a,
                        k_b, l_b, delta_b, eta_b, sigma_b, r_m, r_max,
                        delta_phi=np.pi / 4) -> float:
    """dwba_outer_integral_1d (oscillatory_integrals.py:459-535):
    1/2 [Re I(Phi_-) - Re I(Phi_+)]."""
    if r_max <= r_m + 1e-10:
        return 0.0
    pf = PhaseFunctions(k_a, l_a, delta_a, eta_a, sigma_a, k_b, l_b, delta_b, eta_b, sigma_b)
    im = outer_integral_oscillatory(envelope, lambda r: pf.phi(-1, r), lambda r: pf.dphi(-1, r),
                                    lambda r: pf.d2phi(-1, r), r_m, r_max, delta_phi)
    ip = outer_integral_oscillatory(envelope, lambda r: pf.phi(+1, r), lambda r: pf.dphi(+1, r),
                                    lambda r: pf.d2phi(+1, r), r_m, r_max, delta_phi)
    return 0.5 * (im.real - ip.real)


def analytical_multipole_tail(r_m, k_i, k_f, delta_i, delta_f, l_i, l_f, L,
                              bound_overlap=1.0, eta_i=0.0, eta_f=0.0,
                              sigma_i=0.0, sigma_f=0.0) -> float:
    """_analytical_multipole_tail (oscillatory_integrals.py:2064-2184)."""
    if r_m < 1e-3 or abs(bound_overlap) < 1e-12:
        return 0.0
    lt_i = eta_i * np.log(2 * k_i * r_m + 1e-30) if abs(eta_i) > 1e-10 else 0.0
    lt_f = eta_f * np.log(2 * k_f * r_m + 1e-30) if abs(eta_f) > 1e-10 else 0.0
    k_diff = k_i - k_f
    k_sum = k_i + k_f
    n = L + 1
    tail = 0.0
    pic = -l_i * np.pi / 2 + sigma_i + delta_i + lt_i
    pfc = -l_f * np.pi / 2 + sigma_f + delta_f + lt_f
    phi_diff = k_diff * r_m + pic - pfc
    if abs(k_diff) > 1e-6:
        t1 = np.sin(phi_diff) / (k_diff * (r_m ** n))
        t1c = n * np.cos(phi_diff) / (k_diff ** 2 * (r_m ** (n + 1)))
        tail += 0.5 * (t1 + t1c)
    elif n > 1:
        tail += 0.5 * np.cos(phi_diff) / ((n - 1) * (r_m ** (n - 1)))
    phi_sum = k_sum * r_m + pic + pfc
    if k_sum > 1e-6:
        t2 = np.sin(phi_sum) / (k_sum * (r_m ** n))
        t2c = n * np.cos(phi_sum) / (k_sum ** 2 * (r_m ** (n + 1)))
        tail -= 0.5 * (t2 + t2c)
    if not np.isfinite(tail):
        return 0.0
    return float(tail * bound_overlap)


def outer_integral_noise(envelope, k_a, l_a, delta_a, eta_a, sigma_a, k_b, l_b, delta_b, eta_b, sigma_b,
                         r_m, r_max, delta_phi=np.pi / 4) -> float:
    """Rounding-noise bound of dwba_outer_integral: 0.5 * sum over the Levin
    segments of both phases of cond(A_s) * 2^-53 * |J_s|.  The Levin
    collocation systems become ill-conditioned when |Phi'| is small (the
    Chebyshev matrix has the constants in its null space), so two correct
    solvers (LAPACK zgesv and any other LU) legitimately differ by this much;
    Filon segments contribute nothing at this level."""
    if r_max <= r_m + 1e-10:
        return 0.0
    pf = PhaseFunctions(k_a, l_a, delta_a, eta_a, sigma_a, k_b, l_b, delta_b, eta_b, sigma_b)
    tot = 0.0
    x_ref = np.cos(np.pi * np.arange(8) / 7)[::-1]
    Dref = chebyshev_diff_matrix(8)
    for sg in (-1, 1):
        mid = (r_m + r_max) / 2
        om = abs(pf.dphi(sg, mid))
        if om < 1e-10:
            continue
        h = delta_phi / om
        if abs(pf.d2phi(sg, mid)) * h * h < 0.1:
            continue
        n_seg = max(1, int(np.ceil((r_max - r_m) / h)))
        bounds = np.linspace(r_m, r_max, n_seg + 1)
        for i in range(n_seg):
            nodes = 0.5 * (bounds[i + 1] - bounds[i]) * (x_ref + 1) + bounds[i]
            A = Dref * (2.0 / (nodes[-1] - nodes[0])) + 1j * np.diag(pf.dphi(sg, nodes))
            fv = np.array([envelope(x) for x in nodes], dtype=float)
            J = levin_segment(fv, nodes, pf.phi(sg, nodes), pf.dphi(sg, nodes))
            tot += np.linalg.cond(A) * 2.0 ** -53 * (abs(J) + np.max(np.abs(fv)) * (nodes[-1] - nodes[0]))
    return 0.5 * tot
