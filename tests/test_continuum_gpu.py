"""GPU checks of the batched continuum-wave solver (dwba_amd.continuum).

* Against the reference's own waves (tests/golden/continuum_johnson_c1grid.npz,
  solve_continuum_wave(solver="johnson") on the C1 grid): every case, solved
  in batches of all its l at once; waves to 1e-11 of their maximum, phase
  shifts to 1e-11 rad, match points and solver path equal.  The device
  spherical Bessel functions stand in for scipy's (1e-13 relative).
* Against the oracle on a longer grid and more partial waves (l to 60), with
  both solver chains and the log-derivative phase.
* The reference's default solver ("auto": adaptive RK45 for low l) wrote the
  C1 fixture waves; the grid recursions agree with them to the ODE tolerance
  in the phase shift (1e-4 rad), not to rounding.
"""
import os

import numpy as np
import pytest

from oracle import continuum as OC

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


class _G:
    def __init__(self, r, w):
        self.r, self.w_trapz = r, w


@pytest.fixture(scope="module")
def data():
    g = np.load(os.path.join(GOLD, "continuum_johnson_c1grid.npz"))
    c1 = np.load(os.path.join(GOLD, "c1_h1s2p_50eV.npz"))
    return g, c1


def test_reference_waves_batched(data):
    from dwba_amd.continuum import solve_continuum_waves
    g, c1 = data
    pots = [c1["U_i"], c1["U_f"], c1["V_core"]]
    grid = _G(c1["r"], c1["w_trapz"])
    groups = {}
    for i, case in enumerate(g["cases"]):
        key = (int(case[0]), float(case[1]), float(case[3]), int(case[4]), int(case[5]))
        groups.setdefault(key, []).append(i)
    n = 0
    for (pot, E, z, ph, byp), idx in groups.items():
        ls = [int(g["cases"][i][2]) for i in idx]
        ws = solve_continuum_waves(grid, pots[pot], ls, E, z, ["hybrid", "logderiv"][ph], "johnson", bool(byp))
        for i, w in zip(idx, ws):
            chi, sc, meth = g["chi"][i], g["scalars"][i], str(g["methods"][i])
            assert w is not None
            assert np.abs(w.chi_of_r - chi).max() <= 1e-11 * np.abs(chi).max(), (g["cases"][i], meth)
            assert abs(w.phase_shift - sc[0]) <= 1e-11, (g["cases"][i], meth)
            assert w.idx_match == int(sc[1])
            assert w.solver_method.split("/")[0] == meth.split("|")[0].split("/")[0], (w.solver_method, meth)
            n += 1
    assert n == len(g["cases"])


@pytest.mark.parametrize("solver,phase", [("johnson", "hybrid"), ("numerov", "logderiv")])
def test_many_partial_waves_against_oracle(data, solver, phase):
    from dwba_amd.continuum import solve_continuum_waves
    from dwba_amd.grid import exp_grid
    g, c1 = data
    gr = exp_grid(6000, 1e-4, 300.0)
    r = gr.r
    U = -(1.0 + 1.0 / r) * np.exp(-2.0 * r)
    ls = list(range(0, 61, 6))
    ws = solve_continuum_waves(gr, U, ls, 120.0, 0.0, phase, solver, False)      # every l propagated
    for l, w in zip(ls, ws):
        ref = OC.solve(r, gr.w_trapz, U, l, 120.0, 0.0, phase, False, solver)
        assert np.abs(w.chi_of_r - ref.chi_of_r).max() <= 1e-10 * np.abs(ref.chi_of_r).max(), l
        assert abs(w.phase_shift - ref.phase_shift) <= 1e-10, l
    assert all(w is not None for w in ws)


def test_ionic_waves(data):
    from dwba_amd.continuum import solve_continuum_waves
    g, c1 = data
    r, w = c1["r"], c1["w_trapz"]
    V = 2.0 * c1["V_core"]
    ls = [0, 2, 7, 15]
    ws = solve_continuum_waves(_G(r, w), V, ls, 80.0, 2.0, "hybrid", "johnson", True)
    for l, wv in zip(ls, ws):
        ref = OC.solve(r, w, V, l, 80.0, 2.0, "hybrid", True)
        assert np.abs(wv.chi_of_r - ref.chi_of_r).max() <= 1e-10 * np.abs(ref.chi_of_r).max(), l
        assert abs(wv.phase_shift - ref.phase_shift) <= 1e-10 and abs(wv.sigma_l - ref.sigma_l) <= 1e-13
        assert abs(wv.eta - ref.eta) == 0.0


def test_default_solver_fixture_waves_to_ode_tolerance(data):
    from dwba_amd.continuum import solve_continuum_waves
    g, c1 = data
    r, w = c1["r"], c1["w_trapz"]
    mi = c1["meta_i"]
    ls = [int(x) for x in mi[:, 1]]
    ws = solve_continuum_waves(_G(r, w), c1["U_i"], ls, 50.0, 0.0, "hybrid", "auto", True)
    for q, wv in enumerate(ws):
        d = abs(wv.phase_shift - float(mi[q, 2]))
        assert min(d, abs(d - np.pi)) <= 1e-4, (ls[q], wv.phase_shift, mi[q, 2], wv.solver_method)


def test_refusals():
    from dwba_amd.continuum import solve_continuum_wave
    g = _G(np.linspace(0.1, 50.0, 400), np.full(400, 0.125))
    with pytest.raises(ValueError):
        solve_continuum_wave(g, np.zeros(400), 0, -1.0)
    with pytest.raises(ValueError):
        solve_continuum_wave(g, np.zeros(400), 0, 10.0, solver="rk45")
    with pytest.raises(ValueError):
        solve_continuum_wave(g, np.zeros(300), 0, 10.0)
