"""CPU checks of the amplitude-assembly oracle and of the host-side angular
tables against the reference's own outputs (tests/golden/amplitudes_c1.npz,
written from dwba_coupling.py / sigma_total.py on the C1 integrals).

The tables of dwba_amd.angular are contracted here exactly as the GPU kernels
contract them (a = scale phase sum_L C_L I_L, F = sum_pairs a Y_{l_f, m}), so a
wrong coefficient, phase or m selection fails before any GPU run.
"""
import math
import os

import numpy as np
import pytest

from dwba_amd import angular
from oracle import amplitude as OA

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def gold():
    return np.load(os.path.join(GOLD, "amplitudes_c1.npz"))


@pytest.fixture(scope="module")
def c1():
    return np.load(os.path.join(GOLD, "c1_h1s2p_50eV.npz"))


def test_wigner_symbols_oracle_and_exact(gold):
    for row in gold["w3j"]:
        a, v = [int(x) for x in row[:6]], row[6]
        assert abs(OA.wigner_3j(*a) - v) <= 1e-13 * max(1.0, abs(v))
        assert abs(angular.three_j(*a) - v) <= 1e-13 * max(1.0, abs(v))
    for row in gold["cg"]:
        a, v = [int(x) for x in row[:6]], row[6]
        assert abs(OA.clebsch_gordan(*a) - v) <= 1e-13 * max(1.0, abs(v))
        assert abs(angular.clebsch(*a) - v) <= 1e-13 * max(1.0, abs(v))
    for row in gold["w6j"]:
        a, v = [int(x) for x in row[:6]], row[6]
        assert abs(OA.wigner_6j(*a) - v) <= 1e-13 * max(1.0, abs(v))
        assert abs(angular.six_j(*a) - v) <= 1e-13 * max(1.0, abs(v))
    for row in gold["racah"]:
        a, v = [int(x) for x in row[:6]], row[6]
        assert abs(OA.racah_w(*a) - v) <= 1e-13 * max(1.0, abs(v))
        assert abs(angular.racah(*a) - v) <= 1e-13 * max(1.0, abs(v))


def test_exact_symbols_closed_forms():
    # (j j 0; m -m 0) = (-1)^(j-m)/sqrt(2j+1); sum_m3 of CG^2 over j3 = 1 (orthonormality);
    # {a b c; 0 c b} = (-1)^(a+b+c)/sqrt((2b+1)(2c+1)).
    for j in range(8):
        for m in range(-j, j + 1):
            assert abs(angular.three_j(j, j, 0, m, -m, 0) - (-1) ** (j - m) / math.sqrt(2 * j + 1)) < 1e-15
    for j1, j2, m1, m2 in [(2, 3, 1, -2), (4, 4, 0, 0), (1, 5, -1, 3)]:
        s = sum(angular.clebsch(j1, j2, j3, m1, m2, m1 + m2) ** 2 for j3 in range(abs(j1 - j2), j1 + j2 + 1))
        assert abs(s - 1.0) < 1e-14
    for a, b, c in [(1, 2, 3), (2, 2, 2), (0, 4, 4), (3, 5, 4)]:
        want = (-1) ** (a + b + c) / math.sqrt((2 * b + 1) * (2 * c + 1))
        assert abs(angular.six_j(a, b, c, 0, c, b) - want) < 1e-15


def _contract(rows, I, scale, ylm_d, ylm_x):
    """The kernels' arithmetic in numpy: per projection pair, f and g."""
    out = []
    for (cd, cx, pd, px), yd, yx in zip(rows, ylm_d, ylm_x):
        Id, Ix = I[0].copy(), I[1].copy()
        Id[np.abs(Id) < 1e-20] = 0.0
        Ix[np.abs(Ix) < 1e-20] = 0.0
        out.append((scale * pd * float(np.dot(cd, Id)) * yd, scale * px * float(np.dot(cx, Ix)) * yx))
    return out


def test_oracle_amplitudes_match_reference(gold, c1):
    theta = gold["theta"]
    I = c1["I_full_split_cc5"]
    ki, kf = float(gold["k_i"]), float(gold["k_f"])
    F = np.zeros((3, theta.size), complex)
    G = np.zeros_like(F)
    for q, (li, lf) in enumerate(c1["pairs"]):
        d = {L: float(I[q, 0, L]) for L in range(I.shape[2])}
        e = {L: float(I[q, 1, L]) for L in range(I.shape[2])}
        for c, Mf in enumerate(gold["Mf"]):
            f, g = OA.amplitude_contribution(theta, d, e, int(li), int(lf), ki, kf, 0, 1, 0, int(Mf))
            F[c] += f
            G[c] += g
    scale = max(np.abs(gold["F"]).max(), np.abs(gold["G"]).max())
    assert np.abs(F - gold["F"]).max() <= 1e-12 * scale
    assert np.abs(G - gold["G"]).max() <= 1e-12 * scale
    dcs = sum(OA.dcs_dwba(theta, F[c], G[c], ki, kf, 0, 1) for c in range(3))
    assert np.abs(dcs - gold["dcs"]).max() <= 1e-12 * gold["dcs"].max()
    assert abs(OA.integrate_dcs(theta, gold["dcs"]) - float(gold["tcs"])) <= 1e-14 * float(gold["tcs"])


def test_angular_tables_reproduce_reference(gold, c1):
    """Tables + the kernels' contraction order, Y_lm from scipy, against the golden F, G."""
    theta = gold["theta"]
    I = c1["I_full_split_cc5"]
    L_max = I.shape[2] - 1
    ki, kf = float(gold["k_i"]), float(gold["k_f"])
    combos = [(0, int(m)) for m in gold["Mf"]]
    F = np.zeros((3, theta.size), complex)
    G = np.zeros_like(F)
    for q, (li, lf) in enumerate(c1["pairs"]):
        li, lf = int(li), int(lf)
        rows = angular.excitation_channel(li, lf, 0, 1, combos, L_max)
        yd = [OA.ylm(lf, mi - mf, theta) if abs(mi - mf) <= lf else 0 * theta for mi, mf in combos]
        yx = [OA.ylm(lf, mf - mi, theta) if abs(mf - mi) <= lf else 0 * theta for mi, mf in combos]
        for c, (f, g) in enumerate(_contract(rows, I[q], (2 / math.pi) / (ki * kf), yd, yx)):
            F[c] += f
            G[c] += g
    scale = max(np.abs(gold["F"]).max(), np.abs(gold["G"]).max())
    assert np.abs(F - gold["F"]).max() <= 1e-12 * scale
    assert np.abs(G - gold["G"]).max() <= 1e-12 * scale


def test_ionization_oracle_and_tables(gold):
    for args, Iv, want in zip(gold["ion_args"], gold["ion_I"], gold["ion_out"]):
        li, lf, le, Li, Mi, Mf = (int(x) for x in args[:6])
        ki, kf, ke = (float(x) for x in args[6:])
        d = {L: float(Iv[0][L]) for L in range(Iv.shape[1])}
        e = {L: float(Iv[1][L]) for L in range(Iv.shape[1])}
        f, g = OA.ionization_coefficients(d, e, li, lf, le, ki, kf, ke, Li, Mi, Mf)
        sc = max(1e-300, abs(want[0]), abs(want[1]))
        assert abs(f - want[0]) <= 1e-12 * sc and abs(g - want[1]) <= 1e-12 * sc
        rows = angular.ionization_channel(li, lf, le, Li, [(Mi, Mf)], Iv.shape[1] - 1)
        scale = (2 / math.pi) / (ki * kf) * math.sqrt(2 / math.pi) / ke
        (ft, gt), = _contract(rows, Iv, scale, [1.0], [1.0])
        assert abs(ft - want[0]) <= 1e-12 * sc and abs(gt - want[1]) <= 1e-12 * sc


def test_oracle_mutation_is_caught(gold, c1):
    """A dropped (-1)^nu in the exchange phase must break the golden match."""
    theta = gold["theta_sub"]
    I = c1["I_full_split_cc5"]
    ki, kf = float(gold["k_i"]), float(gold["k_f"])
    q = int(gold["sub_pairs"][3])
    li, lf = (int(x) for x in c1["pairs"][q])
    rows = angular.excitation_channel(li, lf, 0, 1, [(0, -1)], I.shape[2] - 1)
    cd, cx, pd, px = rows[0]
    bad = [(cd, cx, pd, -px)]
    yx = [OA.ylm(lf, -1, theta)]
    (_, g_bad), = _contract(bad, I[q], (2 / math.pi) / (ki * kf), [0 * theta], yx)
    want = gold["g_sub"][3, 0]
    assert np.abs(want).max() > 0
    assert np.abs(g_bad - want).max() > 1e-3 * np.abs(want).max()
