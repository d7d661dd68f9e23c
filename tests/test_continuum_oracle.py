"""CPU checks of the continuum-wave oracle (oracle/continuum.py) against the
reference's own solver output (tests/golden/continuum_johnson_c1grid.npz,
written by make_golden_continuum.py from solve_continuum_wave(solver="johnson")
on the C1 grid and potentials): 92 waves -- neutral at two energies, l up to
45, phase extraction "hybrid" and "logderiv", analytic bypass on and off;
ionic at z = 1 -- covering the Johnson path, its Numerov fallback and the
Born bypass.  The restatement reproduces every wave, phase shift and match
point to rounding.  Closed forms pin the helpers: Fornberg's 5-point weights
are exact for quartics; the Riccati-Bessel pair has Wronskian 1.
"""
import math
import os

import numpy as np
import pytest

from oracle import continuum as OC

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def data():
    g = np.load(os.path.join(GOLD, "continuum_johnson_c1grid.npz"))
    c1 = np.load(os.path.join(GOLD, "c1_h1s2p_50eV.npz"))
    return g, c1


def test_oracle_reproduces_reference_waves(data):
    g, c1 = data
    pots = [c1["U_i"], c1["U_f"], c1["V_core"]]
    seen = set()
    for case, chi, sc, meth in zip(g["cases"], g["chi"], g["scalars"], g["methods"]):
        pot, E, l, z = int(case[0]), float(case[1]), int(case[2]), float(case[3])
        ph, byp = ["hybrid", "logderiv"][int(case[4])], bool(case[5])
        w = OC.solve(c1["r"], c1["w_trapz"], pots[pot], l, E, z, ph, byp)
        assert np.abs(w.chi_of_r - chi).max() <= 1e-12 * np.abs(chi).max(), (case, meth)
        assert abs(w.phase_shift - sc[0]) <= 1e-12, case
        assert w.idx_match == int(sc[1]) and abs(w.k_au - sc[5]) == 0.0
        seen.add(w.method)
    assert {"born", "numerov", "johnson"} <= seen


def test_fornberg_exact_for_quartics():
    r = np.cumsum(np.linspace(0.01, 0.05, 12))
    f = 3.0 - 2.0 * r + 0.5 * r ** 2 + 0.25 * r ** 3 - 0.125 * r ** 4
    df = -2.0 + r + 0.75 * r ** 2 - 0.5 * r ** 3
    for i in range(len(r)):
        assert abs(OC.derivative_5point(f, r, i) - df[i]) <= 1e-9


def test_riccati_wronskian():
    for l in (0, 1, 5, 20):
        for rho in (0.7, 3.0, 25.0, 80.0):
            jh, jp = OC.riccati_jn(l, rho)
            nh, npr = OC.riccati_yn(l, rho)
            assert abs(jh * npr - jp * nh - 1.0) <= 1e-9, (l, rho)
