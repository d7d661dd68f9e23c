"""CPU checks of the partial-wave loop oracle (oracle/partial_waves.py).

Pinned by: the reference's own driver-order sums (tests/golden/amplitudes_c1.npz
holds F, G, DCS and TCS of all 220 C1 pairs, i.e. the excitation loop at
L_max_proj = 20 with waves up to l = 20); a closed case of the stopping rule
(contributions that vanish from l_i = 10 on leave the running DCS unchanged
from there, so the 4-deep history first shows no change at l_i = 12 and the
rule, which also needs l_i > 15, stops at 16); linearity in the set of
partial waves; the identity between the incremental coherent sum and the
spin sum of the final coherent amplitudes.
"""
import math
import os

import numpy as np
import pytest

from oracle import amplitude as OA
from oracle import partial_waves as OP

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def c1():
    c = np.load(os.path.join(GOLD, "c1_h1s2p_50eV.npz"))
    g = np.load(os.path.join(GOLD, "amplitudes_c1.npz"))
    I = c["I_full_split_cc5"]
    idx = {(int(a), int(b)): q for q, (a, b) in enumerate(c["pairs"])}

    def integ(li, lf, cut=None):
        if cut is not None and li >= cut:
            return {L: 0.0 for L in range(5)}, {L: 0.0 for L in range(5)}
        q = idx[(li, lf)]
        return {L: float(I[q, 0, L]) for L in range(5)}, {L: float(I[q, 1, L]) for L in range(5)}
    return g, integ


def test_loop_reproduces_reference_sums(c1):
    g, integ = c1
    r = OP.excitation_loop(g["theta"], integ, lambda l: True, lambda l: l <= 20, float(g["k_i"]), float(g["k_f"]),
                           0, 1, 4, 20)
    assert r["stop_reason"] is None
    sc = np.abs(g["F"]).max()
    assert np.abs(r["F"] - g["F"]).max() <= 1e-13 * sc
    assert np.abs(r["G"] - g["G"]).max() <= 1e-13 * sc
    assert np.abs(r["dcs"] - g["dcs"]).max() <= 1e-12 * g["dcs"].max()
    assert abs(r["sigma_au"] - float(g["tcs"])) <= 1e-12 * float(g["tcs"])
    assert sorted(r["partial_waves"]) == sorted(f"L{l}" for l in range(21))


def test_stop_rule_closed_case(c1):
    g, integ = c1
    r = OP.excitation_loop(g["theta"], lambda li, lf: integ(li, lf, cut=10), lambda l: True, lambda l: l <= 20,
                           float(g["k_i"]), float(g["k_f"]), 0, 1, 4, 30)
    assert r["stop_L"] == 16 and r["stop_reason"] == "Converged"
    part = OP.excitation_loop(g["theta"], integ, lambda l: l < 10, lambda l: l <= 20, float(g["k_i"]),
                              float(g["k_f"]), 0, 1, 4, 20)
    assert np.array_equal(r["F"], part["F"]) and np.array_equal(r["G"], part["G"])
    assert r["partial_waves"]["L12"] == 0.0 and "L17" not in r["partial_waves"]


def test_missing_waves_are_skipped_linearly(c1):
    g, integ = c1
    k = (float(g["k_i"]), float(g["k_f"]))
    without = OP.excitation_loop(g["theta"], integ, lambda l: l != 3, lambda l: l <= 20 and l != 6, *k, 0, 1, 4, 12)
    only3 = OP.excitation_loop(g["theta"], integ, lambda l: l == 3, lambda l: l <= 20 and l != 6, *k, 0, 1, 4, 12)
    both = OP.excitation_loop(g["theta"], integ, lambda l: True, lambda l: l <= 20 and l != 6, *k, 0, 1, 4, 12)
    assert without["skipped"] == 1 and "L3" not in without["partial_waves"]
    sc = np.abs(both["F"]).max()
    assert np.abs(without["F"] + only3["F"] - both["F"]).max() <= 1e-13 * sc
    assert np.abs(without["G"] + only3["G"] - both["G"]).max() <= 1e-13 * sc


def _ion_integrals(seed=3):
    rng = np.random.default_rng(seed)
    cache = {}

    def integ(li, le, lf):
        key = (li, le, lf)
        if key not in cache:
            cache[key] = ({L: float(v) for L, v in enumerate(rng.standard_normal(5) * 0.3 ** (li + lf))},
                          {L: float(v) for L, v in enumerate(rng.standard_normal(5) * 0.3 ** (li + lf))})
        return cache[key]
    return integ


def test_ionization_incremental_coherent_sum():
    integ = _ion_integrals()
    r = OP.ionization_loop(integ, lambda l: True, lambda l: True, lambda l: True, 2.0, 1.1, 0.7, 1, 2, 4, 8, 12)
    pref = (1.1 * 0.7 / 2.0) * (1.0 / 3.0) * (2 * math.pi) ** 4
    direct = pref * sum(0.25 * abs(f + g) ** 2 + 0.75 * abs(f - g) ** 2 for f, g in r["coherent"].values())
    assert abs(r["sigma_au"] - direct) <= 1e-12 * direct
    assert r["last_l_i"] <= 8


def test_ionization_single_wave_is_incoherent():
    integ = _ion_integrals(5)
    r = OP.ionization_loop(integ, lambda l: l == 0, lambda l: True, lambda l: True, 2.0, 1.1, 0.7, 0, 3, 4, 6, 10)
    assert r["last_l_i"] == 0
    assert abs(r["sigma_au"] - r["partial_L"]["L0"]) <= 1e-13 * r["sigma_au"]
