"""Oracle: the continuum-wave solver on its deterministic paths.

Test infrastructure only.  Restates solve_continuum_wave
(continuum.py:1439-2200) for solver="johnson", step by step:

  find_match_point        continuum.py:250-368
  analytic bypass         :1620-1690 (Born phase -k sum w U j_l^2 r^2,
                          Riccati-Bessel or asymptotic Coulomb wave)
  start index             :1702-1716 (0.9 r_turn when inside the barrier)
  johnson                 _johnson_log_derivative_solve, :1180-1332
  derivative_5point       :924-983 (Fornberg weights on the local 5 points)
  phase                   logderiv :370-459, lsq fits :656-767, hybrid :770-921
  stitching, amplitude    :2030-2126 (scale from chi and chi' at r_m, the
                          asymptotic wave beyond, RMS amplitude forced to 1
                          when it is off by more than 5%)

with scipy's spherical Bessel functions and loggamma, as the reference uses.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
from scipy.special import loggamma, spherical_jn, spherical_yn

HARTREE_TO_EV = 27.211386245988
EV_TO_HARTREE = 1.0 / HARTREE_TO_EV


def k_from_E_eV(E_eV):
    """grid.py:105-132: sqrt(2 E_au), E_au = E_eV * (1 / 27.211...)."""
    return float(np.sqrt(2.0 * max(E_eV * EV_TO_HARTREE, 0.0)))


@dataclass
class Wave:
    l: int
    k_au: float
    chi_of_r: np.ndarray
    phase_shift: float
    r_match: float = 0.0
    idx_match: int = -1
    eta: float = 0.0
    sigma_l: float = 0.0
    method: str = ""


def find_match_point(r, U, k, l, threshold=1e-2, idx_start=0, z_ion=0.0, min_search_fraction=0.25):
    k2 = k * k
    N = len(r)
    frac = float(np.clip(min_search_fraction, 0.0, 0.9))
    s0 = max(idx_start + 50, int(frac * N))
    s1 = int(0.9 * N)
    r_turn = math.sqrt(l * (l + 1)) / k if k > 1e-6 else 0.0
    s0 = max(s0, int(np.searchsorted(r, r_turn)) + 10)
    r_safe = 2.5 * r_turn
    r_coul = 3.0 * max(l, abs(z_ion) / k) / k if (abs(z_ion) > 1e-6 and k > 1e-6) else 0.0
    for i in range(s0, s1):
        ri = r[i]
        if ri < r_safe or ri < r_coul:
            continue
        if abs(2.0 * U[i]) + l * (l + 1) / (ri * ri) < threshold * k2:
            return i, ri
    it = int(np.searchsorted(r, r_safe)) if r_safe > 0 else 0
    ic = int(np.searchsorted(r, r_coul)) if r_coul > 0 else 0
    fb = min(max(s0, int(0.7 * N), it + 20, ic + 5), N - 10)
    return fb, r[fb]


def riccati_jn(l, rho):
    if rho < 1e-10:
        return 0.0, 0.0
    jl = float(spherical_jn(l, rho))
    if l == 0:
        return rho * jl, math.cos(rho)
    return rho * jl, rho * float(spherical_jn(l - 1, rho)) - l * jl


def riccati_yn(l, rho):
    if rho < 1e-10:
        return -1e30, -1e30
    yl = float(spherical_yn(l, rho))
    if not np.isfinite(yl) or abs(yl) > 1e50:
        return -1e30, -1e30
    if l == 0:
        return rho * yl, math.sin(rho)
    ym = float(spherical_yn(l - 1, rho))
    if not np.isfinite(ym) or abs(ym) > 1e50:
        return rho * yl, -1e30
    return rho * yl, rho * ym - l * yl


def coulomb_fg(l, eta, rho):
    if rho < 1e-10:
        return 0.0, 0.0, 1e30, 1e30
    sig = float(np.imag(loggamma(l + 1 + 1j * eta)))
    tb = rho + eta * math.log(2.0 * rho) - l * math.pi / 2.0 + sig
    ll1 = l * (l + 1)
    th = tb - (ll1 / (2.0 * rho) if rho > 1.0 else 0.0)
    amp = 1.0 + (ll1 - 2.0 * eta * eta) / (4.0 * rho * rho) if rho > 5.0 else 1.0
    tp = 1.0 + eta / rho + ll1 / (2.0 * rho * rho)
    return amp * math.sin(th), amp * tp * math.cos(th), amp * math.cos(th), -amp * tp * math.sin(th)


def logderiv_phase(Y, k, r_m, l, z_ion):
    rho = k * r_m
    Yr = Y / k
    if abs(z_ion) < 1e-3:
        jh, jp = riccati_jn(l, rho)
        nh, npr = riccati_yn(l, rho)
        num, den = Yr * jh - jp, Yr * nh - npr
    else:
        F, Fp, G, Gp = coulomb_fg(l, -z_ion / k, rho)
        num, den = Yr * F - Fp, Yr * G - Gp
    if abs(den) < 1e-15:
        return math.pi / 2.0 if num > 0 else -math.pi / 2.0
    return math.atan(num / den)


def lsq_phase(r_tail, chi_tail, l, k, z_ion):
    if abs(z_ion) < 1e-3:
        ph = k * r_tail - float(l) * math.pi / 2.0
    else:
        eta = -z_ion / k
        rho = k * r_tail
        ph = rho + eta * np.log(2.0 * rho) - l * math.pi / 2.0 + float(np.imag(loggamma(l + 1 + 1j * eta)))
    M = np.vstack([np.sin(ph), np.cos(ph)]).T
    (a, b), *_ = np.linalg.lstsq(M, chi_tail, rcond=None)
    A = float(np.hypot(a, b))
    if not np.isfinite(A) or A > 1e100:
        return 0.0, 0.0
    return A, float(np.arctan2(b, a))


def derivative_5point(chi, r, idx):
    N = len(chi)
    n = 5 if N >= 5 else N
    s = min(max(idx - 2, 0), N - n)
    x = r[s:s + n]
    x0 = r[idx]
    c = np.zeros((n, 2))
    c[0, 0] = 1.0
    c1 = 1.0
    c4 = x[0] - x0
    for i in range(1, n):
        mn = min(i, 1)
        c2 = 1.0
        c5 = c4
        c4 = x[i] - x0
        for j in range(i):
            c3 = x[i] - x[j]
            c2 *= c3
            if j == i - 1:
                for kk in range(mn, 0, -1):
                    c[i, kk] = c1 * (kk * c[i - 1, kk - 1] - c5 * c[i - 1, kk]) / c2
                c[i, 0] = -c1 * c5 * c[i - 1, 0] / c2
            for kk in range(mn, 0, -1):
                c[j, kk] = (c4 * c[j, kk] - kk * c[j, kk - 1]) / c3
            c[j, 0] = c4 * c[j, 0] / c3
        c1 = c2
    return float(np.dot(c[:, 1], chi[s:s + n]))


def johnson(r, U, l, k):
    N = len(r)
    k2 = k * k
    ell = float(l)
    Q = ell * (ell + 1.0) / (r * r) + 2.0 * U - k2
    R = np.zeros(N)
    r0, r1 = r[0], r[1]
    if Q[0] > 0:
        R[1] = math.exp(-math.sqrt(Q[0]) * (r1 - r0))
    else:
        rho0, rho1 = k * r0, k * r1
        if rho0 > 1e-10 and rho1 > 1e-10:
            c0, c1 = r0 * spherical_jn(l, rho0), r1 * spherical_jn(l, rho1)
            R[1] = c0 / c1 if abs(c1) > 1e-100 else (r0 / r1) ** (l + 1)
        else:
            R[1] = (r0 / r1) ** (l + 1)
    R[1] = min(max(R[1], -1e10), 1e10)
    for i in range(1, N - 1):
        he = math.sqrt((r[i + 1] - r[i]) * (r[i] - r[i - 1]))
        d = 2.0 - he * he * Q[i] - R[i]
        if abs(d) < 1e-12:
            d = math.copysign(1e-12, d) if d != 0 else 1e-12
        R[i + 1] = min(max(1.0 / d, -1e10), 1e10)
    chi = np.zeros(N)
    chi[-1] = 1.0
    for i in range(N - 1, 0, -1):
        chi[i - 1] = R[i] * chi[i]
        if i % 100 == 0:
            mv = np.max(np.abs(chi[i - 1:]))
            if mv > 1e30 or (0 < mv < 1e-30):
                chi[i - 1:] *= 1.0 / mv
    mv = np.max(np.abs(chi))
    if mv > 0:
        chi /= mv
    return chi, np.array([derivative_5point(chi, r, i) for i in range(N)])


def numerov(r, Q, chi0, chi1, renorm_interval=100, renorm_scale=1e50):
    """_numerov_propagate (continuum.py:985-1069): local-step Numerov, every
    renorm_interval steps the values so far rescaled when the last two leave
    [1/renorm_scale, renorm_scale]."""
    N = len(r)
    chi = np.zeros(N)
    chi[0], chi[1] = chi0, chi1
    h = np.diff(r)
    for i in range(1, N - 1):
        h1, h2 = h[i - 1], h[i]
        a_prev = 1.0 - (h1 * h1 / 12.0) * Q[i - 1]
        b_curr = 2.0 + (5.0 * (h1 * h2) / 6.0) * Q[i]
        a_next = 1.0 - (h2 * h2 / 12.0) * Q[i + 1]
        if abs(a_next) < 1e-15:
            a_next = 1e-15 if a_next >= 0 else -1e-15
        chi[i + 1] = (b_curr * chi[i] - a_prev * chi[i - 1]) / a_next
        if (i + 1) % renorm_interval == 0:
            mv = max(abs(chi[i]), abs(chi[i + 1]))
            if mv > renorm_scale:
                chi[:i + 2] *= renorm_scale / mv
            elif 0 < mv < 1.0 / renorm_scale:
                chi[:i + 2] *= 1.0 / (mv * 1e10)
    return chi


def numerov_start(r0, r1, U0, l, k, z_ion):
    """(chi0, chi1) of the Numerov start (continuum.py:1843-1898)."""
    ell = float(l)
    h = r1 - r0
    S0 = ell * (ell + 1.0) / (r0 * r0) + 2.0 * U0 - k * k
    if S0 > 0:
        return 1e-20, 1e-20 * math.exp(math.sqrt(S0) * h)
    if abs(z_ion) >= 1e-3:
        ks = ell * (ell + 1.0) / (r0 * r0) - 2.0 * z_ion / r0 - k * k
        if ks > 0:
            return 1e-20, 1e-20 * math.exp(math.sqrt(ks) * h)
    c0 = r0 * float(spherical_jn(l, k * r0))
    c1 = r1 * float(spherical_jn(l, k * r1))
    nv = max(abs(c0), abs(c1))
    if 0 < nv < 1e-10:
        sc = 1e-10 / nv
        c0, c1 = c0 * sc, c1 * sc
    return c0, c1


def renorm_interval(E_eV):
    """adaptive_renorm_interval of continuum.py:1562-1590."""
    return 50 if E_eV > 5000.0 else (75 if E_eV > 1000.0 else 100)


def asymptotic_wave(r, k, l, delta, z_ion, start=0):
    out = np.zeros(len(r))
    rho = k * r[start:]
    ok = rho >= 1e-10
    rv = rho[ok]
    A = math.sqrt(2.0 / math.pi)
    if abs(z_ion) < 1e-3:
        yl = spherical_yn(l, rv)
        yl = np.where(np.isfinite(yl) & (np.abs(yl) < 1e50), yl, -1e30)
        v = A * (rv * spherical_jn(l, rv) * math.cos(delta) - rv * yl * math.sin(delta))
    else:
        eta = -z_ion / k
        th = rv + eta * np.log(2.0 * rv) - l * math.pi / 2.0 + float(np.imag(loggamma(l + 1 + 1j * eta)))
        v = A * (np.sin(th) * math.cos(delta) - np.cos(th) * math.sin(delta))
    out[np.flatnonzero(ok) + start] = v
    return out


def solve(r, w_trapz, U, l, E_eV, z_ion=0.0, phase="hybrid", bypass=True, solver="johnson"):
    """solve_continuum_wave(grid, U, l, E_eV, z_ion, solver=solver ("johnson":
    Johnson then Numerov; "numerov": the reverse), phase_extraction_method=phase,
    allow_analytic_bypass=bypass)."""
    if E_eV <= 0.0:
        raise ValueError("channel energy must be > 0 eV")
    k = k_from_E_eV(E_eV)
    N = len(r)
    i_chk, r_chk = find_match_point(r, U, k, l, threshold=1e-4, z_ion=z_ion, min_search_fraction=0.0)
    frac = (float(r_chk) - float(r[0])) / max(float(r[-1] - r[0]), 1e-30)
    if bypass and frac < 0.15 and abs(z_ion) <= 1e-6:
        jl = spherical_jn(l, k * r)
        d = float(np.clip(-k * np.sum(w_trapz * (U * jl ** 2 * r ** 2)), -math.pi / 2, math.pi / 2))
        return Wave(l, k, math.sqrt(2.0 / math.pi) * (r * jl * k), d, method="born")
    ell = float(l)
    start = 0
    if ell * (ell + 1.0) / (r[0] * r[0]) + 2.0 * U[0] - k * k > 0:
        rt = math.sqrt(l * (l + 1)) / k if k > 1e-6 else r[-1]
        f = int(np.searchsorted(r, rt * 0.9))
        if 0 < f < N - 20:
            start = f
    # solver order for "johnson": Johnson, then Numerov when its wave vanishes
    # at the match point (continuum.py:1765-1945)
    im, rm = find_match_point(r, U, k, l, threshold=1e-4, idx_start=start, z_ion=z_ion)
    method = None
    for name in (("johnson", "numerov") if solver == "johnson" else ("numerov", "johnson")):
        if name == "johnson":
            co, dco = johnson(r[start:], U[start:], l, k)
        else:
            Q = (ell * (ell + 1.0) / (r * r) + 2.0 * U - k * k)[start:]
            c0, c1 = numerov_start(r[start], r[start + 1], U[start], l, k, z_ion)
            co = numerov(r[start:], Q, c0, c1, renorm_interval(E_eV))
            dco = np.array([derivative_5point(co, r[start:], i) for i in range(N - start)])
        chi = np.zeros(N)
        dchi = np.zeros(N)
        chi[start:] = co
        dchi[start:] = dco
        cm, dm = chi[im], dchi[im]
        if abs(cm) >= 1e-100:
            method = name
            break
    if method is None:
        raise RuntimeError("Johnson and Numerov waves both vanish at the match point")
    if phase == "logderiv":
        delta = logderiv_phase(dm / cm, k, rm, l, z_ion)
    else:
        nt = int(N * 0.15)
        it = max(im, N - nt)
        ok = N - it >= 10 and np.max(np.abs(chi[it:])) > 1e-100
        if ok:
            A, dl = lsq_phase(r[it:], chi[it:], l, k, z_ion)
            ok = np.isfinite(dl) and np.isfinite(A) and A > 1e-100
        if phase == "lsq":
            delta = lsq_phase(r[it:], chi[it:], l, k, z_ion)[1]
        elif ok:
            delta = dl
        else:
            delta = logderiv_phase(dm / cm, k, rm, l, z_ion)
    cas = asymptotic_wave(r, k, l, delta, z_ion)
    dcas = derivative_5point(cas, r, im)
    den = cm * cm + dm * dm
    scale = (cas[im] * cm + dcas * dm) / den if abs(den) > 1e-100 else cas[im] / cm
    fin = np.zeros(N)
    fin[:im + 1] = scale * chi[:im + 1]
    fin[im + 1:] = cas[im + 1:]
    reg, w = fin[im:], w_trapz[im:]
    if len(reg) > 10:
        ws = np.sum(w)
        rms = math.sqrt(np.sum(w * reg ** 2) / ws) if ws > 0 else math.sqrt(np.mean(reg ** 2))
        amp = rms * math.sqrt(2.0)
        if abs(amp - 1.0) > 0.05 and amp > 1e-50:
            fin /= amp
    eta = -z_ion / k if k > 0 else 0.0
    sig = float(np.imag(loggamma(l + 1 + 1j * eta))) if abs(eta) > 1e-10 else 0.0
    return Wave(l, k, fin, delta, rm, im, eta, sig, method=method)
