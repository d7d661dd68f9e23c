"""Oracle: amplitude assembly and cross sections from the radial integrals.

Test infrastructure only.  Restates, step by step:

  wigner_3j / clebsch_gordan / wigner_6j / racah_w   dwba_coupling.py:140-268
  amplitude_contribution   calculate_amplitude_contribution, dwba_coupling.py:285-504
  ionization_coefficients  calculate_ionization_coefficients, dwba_coupling.py:507-634
  dcs_dwba                 sigma_total.py:147-177
  integrate_dcs            integrate_dcs_over_angles, sigma_total.py:179-217

Wigner symbols by the Racah sums in log-factorials, as the reference forms
them (agreement ~1e-14); spherical harmonics by scipy.special.sph_harm
(Condon-Shortley phase), the reference's own routine.
"""
from __future__ import annotations

import math
from functools import lru_cache

import numpy as np


@lru_cache(maxsize=None)
def _lf(n: int) -> float:
    return math.lgamma(n + 1) if n >= 0 else float("-inf")


def _tri(a, b, c) -> float:
    return _lf(a + b - c) + _lf(a - b + c) + _lf(-a + b + c) - _lf(a + b + c + 1)


@lru_cache(maxsize=None)
def wigner_3j(j1, j2, j3, m1, m2, m3) -> float:
    """(j1 j2 j3; m1 m2 m3), integer arguments (dwba_coupling.py:153-185)."""
    if m1 + m2 + m3 != 0 or not (abs(j1 - j2) <= j3 <= j1 + j2):
        return 0.0
    if abs(m1) > j1 or abs(m2) > j2 or abs(m3) > j3:
        return 0.0
    t0 = max(0, j2 - m1 - j3, j1 + m2 - j3)
    t1 = min(j1 + j2 - j3, j1 - m1, j2 + m2)
    if t0 > t1:
        return 0.0
    pre = 0.5 * _tri(j1, j2, j3) + 0.5 * (_lf(j1 + m1) + _lf(j1 - m1) + _lf(j2 + m2) + _lf(j2 - m2) +
                                          _lf(j3 + m3) + _lf(j3 - m3))
    s = 0.0
    for t in range(t0, t1 + 1):
        den = (_lf(t) + _lf(t - (j2 - m1 - j3)) + _lf(t - (j1 + m2 - j3)) + _lf(j1 + j2 - j3 - t) +
               _lf(j1 - m1 - t) + _lf(j2 + m2 - t))
        s += (-1) ** t * math.exp(pre - den)
    return (-1.0) ** (j1 - j2 - m3) * s


def clebsch_gordan(j1, j2, j3, m1, m2, m3) -> float:
    """<j1 m1; j2 m2 | j3 m3> (dwba_coupling.py:187-196)."""
    return (-1.0) ** (j1 - j2 + m3) * math.sqrt(2 * j3 + 1) * wigner_3j(j1, j2, j3, m1, m2, -m3)


@lru_cache(maxsize=None)
def wigner_6j(j1, j2, j3, j4, j5, j6) -> float:
    """{j1 j2 j3; j4 j5 j6} (dwba_coupling.py:198-225)."""
    if not (abs(j1 - j2) <= j3 <= j1 + j2 and abs(j1 - j5) <= j6 <= j1 + j5 and
            abs(j4 - j2) <= j6 <= j4 + j2 and abs(j4 - j5) <= j3 <= j4 + j5):
        return 0.0
    a = (j1 + j2 + j3, j1 + j5 + j6, j4 + j2 + j6, j4 + j5 + j3)
    b = (j1 + j2 + j4 + j5, j2 + j3 + j5 + j6, j3 + j1 + j6 + j4)
    t0, t1 = max(a), min(b)
    if t0 > t1:
        return 0.0
    tri = _tri(j1, j2, j3) + _tri(j1, j5, j6) + _tri(j4, j2, j6) + _tri(j4, j5, j3)
    s = 0.0
    for t in range(t0, t1 + 1):
        den = sum(_lf(t - x) for x in a) + sum(_lf(y - t) for y in b)
        s += (-1) ** t * math.exp(_lf(t + 1) - den)
    return math.exp(0.5 * tri) * s


def racah_w(l1, l2, l3, l4, l5, l6) -> float:
    """W(l1 l2 l3 l4; l5 l6) = (-1)^(l1+l2+l3+l4) {l1 l2 l5; l4 l3 l6} (dwba_coupling.py:227-240)."""
    return (-1.0) ** (l1 + l2 + l3 + l4) * wigner_6j(l1, l2, l5, l4, l3, l6)


def ylm(l, m, theta):
    """Y_lm(theta, phi = 0) with the Condon-Shortley phase (scipy convention;
    sph_harm_y where scipy has it, the same function with the arguments in
    (l, m, theta, phi) order)."""
    import scipy.special as sp
    th = np.asarray(theta, dtype=float)
    if hasattr(sp, "sph_harm_y"):
        return sp.sph_harm_y(l, m, th, np.zeros_like(th))
    return sp.sph_harm(m, l, np.zeros_like(th), th)


def _dir_sum(I_dir, l_i, l_f, Li, Lf, Mi, Mf, mu, w_cut):
    s = 0.0
    for lT, R in I_dir.items():
        if abs(R) < 1e-20:
            continue
        pre = math.sqrt(2 * l_f + 1) * math.sqrt(2 * l_i + 1) * math.sqrt(2 * Li + 1) / math.sqrt(2 * lT + 1)
        c1 = clebsch_gordan(l_f, l_i, lT, 0, 0, 0)
        if abs(c1) < 1e-9:
            continue
        c2 = clebsch_gordan(lT, Li, Lf, 0, 0, 0)
        if abs(c2) < 1e-9:
            continue
        sg = 0.0
        for g in range(max(abs(l_i - Li), abs(l_f - Lf)), min(l_i + Li, l_f + Lf) + 1):
            w = racah_w(l_f, l_i, Lf, Li, lT, g)
            if abs(w) < 1e-9:
                continue
            sg += w * clebsch_gordan(l_i, Li, g, 0, Mi, Mi) * clebsch_gordan(l_f, Lf, g, mu, -Mf, -Mi)
        s += pre * c1 * c2 * sg * R
    return s


def _ex_sum(I_ex, l_i, l_f, Li, Lf, Mi, Mf, mu, w_cut):
    s = 0.0
    for lT, R in I_ex.items():
        if abs(R) < 1e-20:
            continue
        pre = math.sqrt(2 * Lf + 1) * math.sqrt(2 * l_i + 1) * math.sqrt(2 * Li + 1) / math.sqrt(2 * lT + 1)
        c1 = clebsch_gordan(Lf, l_i, lT, 0, 0, 0)
        c2 = clebsch_gordan(lT, Li, l_f, 0, 0, 0)
        if abs(c1 * c2) < 1e-9:
            continue
        sg = 0.0
        for g in range(max(abs(l_i - Li), abs(Lf - l_f)), min(l_i + Li, Lf + l_f) + 1):
            w = racah_w(Lf, l_i, l_f, Li, lT, g)
            if w_cut and abs(w) < 1e-9:
                continue
            sg += w * clebsch_gordan(l_i, Li, g, 0, Mi, Mi) * clebsch_gordan(Lf, l_f, g, -Mf, -mu, -Mi)
        s += pre * c1 * c2 * sg * R
    return s


def amplitude_contribution(theta, I_dir, I_ex, l_i, l_f, ki, kf, Li, Lf, Mi, Mf):
    """(f(theta), g(theta)) of one (l_i, l_f) channel (dwba_coupling.py:285-504):
    direct mu_f = Mf - Mi with Y_{lf,-mu_f}, exchange mu_f = Mi - Mf with
    (-1)^mu_f Y_{lf,-mu_f}; no |W| cut in the exchange g-sum."""
    mu_d, mu_x = Mf - Mi, Mi - Mf
    z = np.zeros_like(theta, dtype=complex)
    if abs(mu_d) > l_f and abs(mu_x) > l_f:
        return z, z.copy()
    y_li = math.sqrt((2 * l_i + 1) / (4 * math.pi))
    pref = (2.0 / math.pi) / (ki * kf)
    y_d = ylm(l_f, -mu_d, theta) if abs(mu_d) <= l_f else z
    f = pref * (1j ** (l_i + l_f)) * y_d * y_li * _dir_sum(I_dir, l_i, l_f, Li, Lf, Mi, Mf, mu_d, True)
    y_x = ((-1.0) ** mu_x) * (ylm(l_f, -mu_x, theta) if abs(mu_x) <= l_f else z)
    g = (pref * (1j ** (l_i - l_f)) * ((-1.0) ** (Lf + Mf)) * y_x * y_li *
         _ex_sum(I_ex, l_i, l_f, Li, Lf, Mi, Mf, mu_x, False))
    return f, g


def ionization_coefficients(I_dir, I_ex, l_i, l_f, l_ej, ki, kf, kej, Li, Mi, Mf, include_eject_norm=True):
    """(f_coeff, g_coeff) of calculate_ionization_coefficients
    (dwba_coupling.py:507-634): the excitation algebra with L_f = l_ej, the
    ejected-wave phase i^-l_ej, the |W| cut in both g-sums, no Y_lm."""
    if ki <= 0.0 or kf <= 0.0 or kej <= 0.0:
        return 0j, 0j
    y_li = math.sqrt((2 * l_i + 1) / (4 * math.pi))
    pref = (2.0 / math.pi) / (ki * kf)
    if include_eject_norm:
        pref *= math.sqrt(2.0 / math.pi) / kej
    ph_ej = 1j ** (-l_ej)
    mu_d, mu_x = Mf - Mi, Mi - Mf
    f = g = 0j
    if abs(mu_d) <= l_f:
        f = pref * (1j ** (l_i + l_f)) * ph_ej * y_li * _dir_sum(I_dir, l_i, l_f, Li, l_ej, Mi, Mf, mu_d, True)
    if abs(mu_x) <= l_f:
        g = (pref * (1j ** (l_i - l_f)) * ((-1.0) ** (l_ej + Mf)) * ph_ej * y_li *
             _ex_sum(I_ex, l_i, l_f, Li, l_ej, Mi, Mf, mu_x, True))
    return f, g


def dcs_dwba(theta, f, g, ki, kf, Li, N_equiv):
    """(k_f/k_i) N_eq/(2 L_i + 1) [1/4 |f+g|^2 + 3/4 |f-g|^2] (2 pi)^4, clipped at 0 (sigma_total.py:147-177)."""
    if ki <= 0.0 or kf <= 0.0:
        return np.zeros_like(theta, dtype=float)
    d = (kf / ki) * (float(N_equiv) / float(2 * Li + 1)) * (0.75 * np.abs(f - g) ** 2 + 0.25 * np.abs(f + g) ** 2)
    return np.clip(np.real(d * (2.0 * np.pi) ** 4), 0.0, None)


def integrate_dcs(theta, dcs):
    """2 pi * trapezoid(sin(theta) dcs, theta) (sigma_total.py:179-217)."""
    y = np.sin(theta) * dcs
    return 2.0 * np.pi * float(np.sum(0.5 * (y[1:] + y[:-1]) * np.diff(theta)))
