"""Write the golden vectors of the reference's continuum-wave solver on its
deterministic (Johnson) path (run once, in the container that has the
reference checked out; the .npz is committed).

Nothing else pins solve_continuum_wave's Johnson propagation, match point,
phase extraction and stitching, so this script imports the reference's
continuum module and records, on the C1 grid (3000 points on [1e-4, 200] a0,
c1_h1s2p_50eV.npz, potentials U_i, U_f and V_core from the reference's own
setup), for solver="johnson":

  neutral, U_i at 50 eV and U_f at 39.8 eV, l in L_SET, phase extraction
      "hybrid" and "logderiv", analytic bypass off and on;
  ionic (z_ion = 1, potential V_core), 50 eV, l in L_SET[:6], "hybrid" and
      "logderiv", bypass on (it is never taken for ions);
each case: chi_of_r, phase_shift, idx_match, r_match, eta, sigma_l and the
solver and phase method the reference reports (Johnson falls back to Numerov,
then RK45, when its wave vanishes at the match point).

Usage:  python tests/golden/make_golden_continuum.py [/root/reference]
"""
import os
import sys
from types import SimpleNamespace

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
L_SET = [0, 1, 2, 3, 5, 8, 12, 20, 30, 45]


def main(ref_root):
    sys.path.insert(0, ref_root)
    os.environ.setdefault("DWBA_LOG_LEVEL", "ERROR")
    import warnings
    warnings.filterwarnings("ignore")
    import continuum as C

    c1 = np.load(os.path.join(HERE, "c1_h1s2p_50eV.npz"))
    grid = SimpleNamespace(r=np.array(c1["r"]), w_trapz=np.array(c1["w_trapz"]), w_simpson=np.array(c1["w_simpson"]))
    cases, chis, scal, meth = [], [], [], []
    pots = {"U_i": np.array(c1["U_i"]), "U_f": np.array(c1["U_f"]), "V_core": np.array(c1["V_core"])}
    specs = []
    for pot, E in (("U_i", 50.0), ("U_f", 39.8)):
        for l in L_SET:
            for ph in ("hybrid", "logderiv"):
                for byp in (False, True):
                    specs.append((pot, E, l, 0.0, ph, byp))
    for l in L_SET[:6]:
        for ph in ("hybrid", "logderiv"):
            specs.append(("V_core", 50.0, l, 1.0, ph, True))
    for (pot, E, l, z, ph, byp) in specs:
        w = C.solve_continuum_wave(grid, SimpleNamespace(U_of_r=pots[pot]), l, E, z_ion=z, phase_extraction_method=ph,
                                   solver="johnson", allow_analytic_bypass=byp)
        cases.append([["U_i", "U_f", "V_core"].index(pot), E, l, z, ["hybrid", "logderiv"].index(ph), int(byp)])
        chis.append(w.chi_of_r)
        scal.append([w.phase_shift, w.idx_match, w.r_match, w.eta, w.sigma_l, w.k_au])
        meth.append(f"{w.solver_method}|{w.phase_method}")
    np.savez_compressed(os.path.join(HERE, "continuum_johnson_c1grid.npz"), cases=np.array(cases),
                        chi=np.array(chis), scalars=np.array(scal), methods=np.array(meth))
    print(len(cases), "cases")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
