"""Angular coupling coefficients of the DWBA amplitudes (host side).

The amplitude of a channel pair is a fixed linear combination of its radial
integrals: for target projections (M_i, M_f),

    f = (2/pi)/(k_i k_f) i^(l_i+l_f) Y*_{l_i 0} Y_{l_f,-mu}(theta) sum_L Cd_L I_dir(L),   mu = M_f - M_i
    g = (2/pi)/(k_i k_f) i^(l_i-l_f) (-1)^(L_f+M_f) Y*_{l_i 0} (-1)^nu Y_{l_f,-nu}(theta) sum_L Cx_L I_ex(L),
                                                                                           nu = M_i - M_f
(calculate_amplitude_contribution, dwba_coupling.py:285-504; the ionization
coefficients, :507-634, are the same sums with L_f = l_ej, an extra phase
i^-l_ej and normalisation, and no Y_lm).  The coefficients

    Cd_L = sqrt((2l_f+1)(2l_i+1)(2L_i+1)/(2L+1)) C(l_f l_i L;000) C(L L_i L_f;000)
           sum_g W(l_f l_i L_f L_i; L g) C(l_i L_i g; 0 M_i M_i) C(l_f L_f g; mu,-M_f,-M_i)
    Cx_L = sqrt((2L_f+1)(2l_i+1)(2L_i+1)/(2L+1)) C(L_f l_i L;000) C(L L_i l_f;000)
           sum_g W(L_f l_i l_f L_i; L g) C(l_i L_i g; 0 M_i M_i) C(L_f l_f g; -M_f,-nu,-M_i)

depend only on angular momenta, so they are formed once per (l_i, l_f,
target, projections) on the host, exactly (Racah's sums in rational
arithmetic, one square root at the end), with the reference's cuts (a
Clebsch-Gordan factor below 1e-9 drops the term; a Racah W below 1e-9 drops
the g term of the direct sum, and of the exchange sum for ionization).  The
sums over L with the integrals run on the GPU (csrc/amplitude.hip).
"""
from __future__ import annotations

import math
from fractions import Fraction
from functools import lru_cache

import numpy as np


@lru_cache(maxsize=None)
def _fact(n: int) -> int:
    return math.factorial(n)


def _tri_sq(a, b, c) -> Fraction:
    """Delta(abc)^2 = (a+b-c)!(a-b+c)!(-a+b+c)!/(a+b+c+1)!."""
    return Fraction(_fact(a + b - c) * _fact(a - b + c) * _fact(-a + b + c), _fact(a + b + c + 1))


@lru_cache(maxsize=None)
def three_j(j1: int, j2: int, j3: int, m1: int, m2: int, m3: int) -> float:
    """Wigner 3j symbol (j1 j2 j3; m1 m2 m3), integer arguments, by Racah's
    formula in exact arithmetic."""
    if m1 + m2 + m3 != 0 or not (abs(j1 - j2) <= j3 <= j1 + j2):
        return 0.0
    if abs(m1) > j1 or abs(m2) > j2 or abs(m3) > j3:
        return 0.0
    lo = max(0, j2 - m1 - j3, j1 + m2 - j3)
    hi = min(j1 + j2 - j3, j1 - m1, j2 + m2)
    if lo > hi:
        return 0.0
    s = Fraction(0)
    for t in range(lo, hi + 1):
        den = (_fact(t) * _fact(t - (j2 - m1 - j3)) * _fact(t - (j1 + m2 - j3)) * _fact(j1 + j2 - j3 - t) *
               _fact(j1 - m1 - t) * _fact(j2 + m2 - t))
        s += Fraction((-1) ** t, den)
    if s == 0:
        return 0.0
    sq = _tri_sq(j1, j2, j3) * (_fact(j1 + m1) * _fact(j1 - m1) * _fact(j2 + m2) * _fact(j2 - m2) *
                                _fact(j3 + m3) * _fact(j3 - m3))
    val = s * s * sq                                   # the square of the symbol, exactly
    sign = (-1) ** (j1 - j2 - m3) * (1 if s > 0 else -1)
    return sign * math.sqrt(val.numerator) / math.sqrt(val.denominator)


def clebsch(j1, j2, j3, m1, m2, m3) -> float:
    """<j1 m1; j2 m2 | j3 m3> = (-1)^(j1-j2+m3) sqrt(2 j3 + 1) (j1 j2 j3; m1 m2 -m3)."""
    return (-1.0) ** (j1 - j2 + m3) * math.sqrt(2 * j3 + 1) * three_j(j1, j2, j3, m1, m2, -m3)


@lru_cache(maxsize=None)
def six_j(j1, j2, j3, j4, j5, j6) -> float:
    """Wigner 6j symbol {j1 j2 j3; j4 j5 j6}, Racah's formula in exact arithmetic."""
    if not (abs(j1 - j2) <= j3 <= j1 + j2 and abs(j1 - j5) <= j6 <= j1 + j5 and
            abs(j4 - j2) <= j6 <= j4 + j2 and abs(j4 - j5) <= j3 <= j4 + j5):
        return 0.0
    a = (j1 + j2 + j3, j1 + j5 + j6, j4 + j2 + j6, j4 + j5 + j3)
    b = (j1 + j2 + j4 + j5, j2 + j3 + j5 + j6, j3 + j1 + j6 + j4)
    lo, hi = max(a), min(b)
    if lo > hi:
        return 0.0
    s = Fraction(0)
    for t in range(lo, hi + 1):
        den = 1
        for x in a:
            den *= _fact(t - x)
        for y in b:
            den *= _fact(y - t)
        s += Fraction((-1) ** t * _fact(t + 1), den)
    if s == 0:
        return 0.0
    val = s * s * _tri_sq(j1, j2, j3) * _tri_sq(j1, j5, j6) * _tri_sq(j4, j2, j6) * _tri_sq(j4, j5, j3)
    return (1 if s > 0 else -1) * math.sqrt(val.numerator) / math.sqrt(val.denominator)


def racah(l1, l2, l3, l4, l5, l6) -> float:
    """Racah W(l1 l2 l3 l4; l5 l6) = (-1)^(l1+l2+l3+l4) {l1 l2 l5; l4 l3 l6}."""
    return (-1.0) ** (l1 + l2 + l3 + l4) * six_j(l1, l2, l5, l4, l3, l6)


@lru_cache(maxsize=None)
def direct_coeffs(l_i, l_f, Li, Lf, Mi, Mf, L_max) -> tuple:
    """Cd_L, L = 0..L_max (see the module docstring)."""
    mu = Mf - Mi
    out = []
    for L in range(L_max + 1):
        c1 = clebsch(l_f, l_i, L, 0, 0, 0)
        c2 = clebsch(L, Li, Lf, 0, 0, 0) if abs(c1) >= 1e-9 else 0.0
        if abs(c1) < 1e-9 or abs(c2) < 1e-9:
            out.append(0.0)
            continue
        sg = 0.0
        for g in range(max(abs(l_i - Li), abs(l_f - Lf)), min(l_i + Li, l_f + Lf) + 1):
            w = racah(l_f, l_i, Lf, Li, L, g)
            if abs(w) < 1e-9:
                continue
            sg += w * clebsch(l_i, Li, g, 0, Mi, Mi) * clebsch(l_f, Lf, g, mu, -Mf, -Mi)
        pre = math.sqrt(2 * l_f + 1) * math.sqrt(2 * l_i + 1) * math.sqrt(2 * Li + 1) / math.sqrt(2 * L + 1)
        out.append(pre * c1 * c2 * sg)
    return tuple(out)


@lru_cache(maxsize=None)
def exchange_coeffs(l_i, l_f, Li, Lf, Mi, Mf, L_max, w_cut) -> tuple:
    """Cx_L, L = 0..L_max; w_cut: drop g terms with |W| < 1e-9 (ionization)."""
    nu = Mi - Mf
    out = []
    for L in range(L_max + 1):
        c1 = clebsch(Lf, l_i, L, 0, 0, 0)
        c2 = clebsch(L, Li, l_f, 0, 0, 0)
        if abs(c1 * c2) < 1e-9:
            out.append(0.0)
            continue
        sg = 0.0
        for g in range(max(abs(l_i - Li), abs(Lf - l_f)), min(l_i + Li, Lf + l_f) + 1):
            w = racah(Lf, l_i, l_f, Li, L, g)
            if w_cut and abs(w) < 1e-9:
                continue
            sg += w * clebsch(l_i, Li, g, 0, Mi, Mi) * clebsch(Lf, l_f, g, -Mf, -nu, -Mi)
        pre = math.sqrt(2 * Lf + 1) * math.sqrt(2 * l_i + 1) * math.sqrt(2 * Li + 1) / math.sqrt(2 * L + 1)
        out.append(pre * c1 * c2 * sg)
    return tuple(out)


def _ipow(n: int) -> complex:
    """i^n exactly."""
    return (1, 1j, -1, -1j)[n % 4]


def excitation_channel(l_i, l_f, Li, Lf, combos, L_max):
    """Per projection pair (Mi, Mf) of `combos`: (Cd, Cx, phase_d, phase_x)
    with the constant complex factors phase_d = i^(l_i+l_f) Y*_{l_i 0} and
    phase_x = i^(l_i-l_f) (-1)^(Lf+Mf) (-1)^nu Y*_{l_i 0}; a family whose
    Y_{l_f,m} does not exist (|m| > l_f) gets phase 0 (and both do when
    neither exists, the reference's early return)."""
    y_li = math.sqrt((2 * l_i + 1) / (4 * math.pi))
    rows = []
    for (Mi, Mf) in combos:
        mu, nu = Mf - Mi, Mi - Mf
        if abs(mu) > l_f and abs(nu) > l_f:
            rows.append((np.zeros(L_max + 1), np.zeros(L_max + 1), 0j, 0j))
            continue
        cd = np.asarray(direct_coeffs(l_i, l_f, Li, Lf, Mi, Mf, L_max))
        cx = np.asarray(exchange_coeffs(l_i, l_f, Li, Lf, Mi, Mf, L_max, False))
        pd = _ipow(l_i + l_f) * y_li if abs(mu) <= l_f else 0j
        px = _ipow(l_i - l_f) * ((-1.0) ** (Lf + Mf)) * ((-1.0) ** nu) * y_li if abs(nu) <= l_f else 0j
        rows.append((cd, cx, complex(pd), complex(px)))
    return rows


def ionization_channel(l_i, l_f, l_ej, Li, combos, L_max):
    """The same for calculate_ionization_coefficients: L_f = l_ej, phases
    i^(l_i+l_f) i^-l_ej Y*_{l_i 0} and i^(l_i-l_f) (-1)^(l_ej+Mf) i^-l_ej Y*_{l_i 0},
    |W| cut in both sums, no Y_lm (so no (-1)^nu)."""
    y_li = math.sqrt((2 * l_i + 1) / (4 * math.pi))
    ph_ej = _ipow(-l_ej)
    rows = []
    for (Mi, Mf) in combos:
        mu, nu = Mf - Mi, Mi - Mf
        cd = np.asarray(direct_coeffs(l_i, l_f, Li, l_ej, Mi, Mf, L_max)) if abs(mu) <= l_f else np.zeros(L_max + 1)
        cx = (np.asarray(exchange_coeffs(l_i, l_f, Li, l_ej, Mi, Mf, L_max, True)) if abs(nu) <= l_f
              else np.zeros(L_max + 1))
        pd = _ipow(l_i + l_f) * ph_ej * y_li if abs(mu) <= l_f else 0j
        px = _ipow(l_i - l_f) * ((-1.0) ** (l_ej + Mf)) * ph_ej * y_li if abs(nu) <= l_f else 0j
        rows.append((cd, cx, complex(pd), complex(px)))
    return rows
