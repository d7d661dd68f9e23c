"""Write the golden vectors of the reference's CPU path (run once, in the
container that has the reference checked out; the .npz files are committed).

The reference's own tests hold no I_L values, and the composite
radial_ME_all_L result on real inputs -- split index, Filon/exchange node
quadrature, outer tails with their Levin branch -- is pinned by nothing
else, so this script imports the reference's CPU modules and records:

  c1_h1s2p_50eV.npz   H 1s -> 2p at 50 eV, Z = 1 Coulomb SAE core, make_r_grid(1e-4, 200, 3000);
                      inputs (grid, V_core, U_i, U_f, u_1s, u_2p, chi_i / chi_f for l = 0..20
                      with k, delta, idx_match, eta, sigma) and radial_ME_all_L outputs for the
                      220 pairs with l_i + l_f odd, L = 0..4 (full_split, CC_nodes = 5), plus
                      subsets with CC_nodes = 7 and the advanced / legacy methods.
  tail_vectors.npz    dwba_outer_integral_1d + _analytical_multipole_tail for parameter tuples
                      covering the Filon, Levin, Coulomb and segment-capped branches, and
                      generate_phase_nodes with Coulomb stepping.

Usage:  python tests/golden/make_golden.py [/root/reference]
"""
from __future__ import annotations

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main(ref_root: str) -> None:
    sys.path.insert(0, ref_root)
    os.environ.setdefault("DWBA_LOG_LEVEL", "ERROR")
    import grid as G
    import potential_core as PC
    import bound_states as BS
    import distorting_potential as DP
    import continuum as CO
    import dwba_matrix_elements as ME
    import oscillatory_integrals as OI

    t0 = time.time()
    g = G.make_r_grid(r_min=1e-4, r_max=200.0, n_points=3000)
    core = PC.CorePotentialParams(Zc=1.0, a1=0.0, a2=0.0, a3=0.0, a4=0.0, a5=0.0, a6=0.0)
    V = PC.V_core_on_grid(g, core)
    s_states = BS.solve_bound_states(g, V, l=0, n_states_max=2)
    p_states = BS.solve_bound_states(g, V, l=1, n_states_max=2)
    orb_i = [o for o in s_states if o.n_index == 1][0]
    orb_f = [o for o in p_states if o.n_index == 1][0]
    E_inc = 50.0
    dE = (orb_f.energy_au - orb_i.energy_au) / G.ev_to_au(1.0)
    E_fin = E_inc - dE
    k_i = float(G.k_from_E_eV(E_inc))
    k_f = float(G.k_from_E_eV(E_fin))
    U_i, U_f = DP.build_distorting_potentials(g, V, orb_i, orb_f, k_i_au=k_i, k_f_au=k_f)
    print(f"setup {time.time() - t0:.1f}s  k_i={k_i:.4f} k_f={k_f:.4f}", flush=True)
    lmax = 20
    chi_i = [CO.solve_continuum_wave(g, U_i, l, E_inc, 0.0) for l in range(lmax + 1)]
    chi_f = [CO.solve_continuum_wave(g, U_f, l, E_fin, 0.0) for l in range(lmax + 1)]
    print(f"waves {time.time() - t0:.1f}s", flush=True)

    def wave_meta(w):
        return [float(w.k_au), float(w.l), float(w.phase_shift), float(w.idx_match), float(getattr(w, "eta", 0.0)),
                float(getattr(w, "sigma_l", 0.0))]

    pairs = [(li, lf) for li in range(lmax + 1) for lf in range(lmax + 1) if (li + lf) % 2 == 1]
    L_max = 4
    out_fs = np.zeros((len(pairs), 2, L_max + 1))
    for q, (li, lf) in enumerate(pairs):
        res = ME.radial_ME_all_L(g, V, U_i.U_of_r, orb_i, orb_f, chi_i[li], chi_f[lf], L_max,
                                 use_oscillatory_quadrature=True, oscillatory_method="full_split", CC_nodes=5,
                                 phase_increment=1.5708, min_grid_fraction=0.1, k_threshold=0.5,
                                 U_f_array=U_f.U_of_r)
        out_fs[q, 0] = [res.I_L_direct[L] for L in range(L_max + 1)]
        out_fs[q, 1] = [res.I_L_exchange[L] for L in range(L_max + 1)]
    print(f"full_split CC5 {time.time() - t0:.1f}s", flush=True)
    sub = list(range(0, len(pairs), 11))       # 20 pairs
    variants = [("full_split", 7), ("advanced", 5), ("legacy", 5)]
    out_var = np.zeros((len(variants), len(sub), 2, L_max + 1))
    for v, (method, cc) in enumerate(variants):
        for s_i, q in enumerate(sub):
            li, lf = pairs[q]
            res = ME.radial_ME_all_L(g, V, U_i.U_of_r, orb_i, orb_f, chi_i[li], chi_f[lf], L_max,
                                     use_oscillatory_quadrature=True, oscillatory_method=method, CC_nodes=cc,
                                     U_f_array=U_f.U_of_r)
            out_var[v, s_i, 0] = [res.I_L_direct[L] for L in range(L_max + 1)]
            out_var[v, s_i, 1] = [res.I_L_exchange[L] for L in range(L_max + 1)]
    print(f"variants {time.time() - t0:.1f}s", flush=True)
    np.savez_compressed(
        os.path.join(HERE, "c1_h1s2p_50eV.npz"),
        r=g.r, w_simpson=g.w_simpson, w_trapz=g.w_trapz, V_core=V, U_i=U_i.U_of_r, U_f=U_f.U_of_r,
        u_i=orb_i.u_of_r, u_f=orb_f.u_of_r, e_i=orb_i.energy_au, e_f=orb_f.energy_au,
        chi_i=np.array([w.chi_of_r for w in chi_i]), chi_f=np.array([w.chi_of_r for w in chi_f]),
        meta_i=np.array([wave_meta(w) for w in chi_i]), meta_f=np.array([wave_meta(w) for w in chi_f]),
        pairs=np.array(pairs), L_max=L_max, I_full_split_cc5=out_fs,
        variant_methods=np.array([m for m, _ in variants]), variant_cc=np.array([c for _, c in variants]),
        variant_pairs=np.array(sub), I_variants=out_var)

    # ---- outer tails and phase nodes
    rng = np.random.default_rng(2024)
    tuples, vals, analytic = [], [], []
    cases = [
        # (k_a, l_a, k_b, l_b, eta_a, eta_b, r_m, r_max, L)
        (2.1, 3, 1.3, 4, 0.0, 0.0, 15.0, 200.0, 1),
        (2.1, 3, 1.3, 4, 0.0, 0.0, 15.0, 200.0, 3),
        (0.5, 10, 0.5, 5, 0.0, 0.0, 5.0, 55.0, 1),       # Levin on Phi_-
        (0.5, 10, 0.5, 5, 0.0, 0.0, 5.0, 55.0, 4),
        (1.7, 2, 1.1, 5, -0.6, -0.9, 20.0, 300.0, 2),    # Coulomb phases
        (9.0, 1, 8.2, 2, 0.0, 0.0, 10.0, 1000.0, 1),     # segment cap + node up-scaling
        (9.0, 1, 8.2, 2, 0.0, 0.0, 10.0, 1000.0, 5),
        (3.0, 0, 0.4, 9, 0.0, 0.0, 8.0, 150.0, 2),
    ]
    for _ in range(12):
        cases.append((rng.uniform(0.3, 6), int(rng.integers(0, 25)), rng.uniform(0.3, 6), int(rng.integers(0, 25)),
                      0.0, 0.0, rng.uniform(5, 60), rng.uniform(100, 1000), int(rng.integers(1, 8))))
    for (ka, la, kb, lb, ea, eb, rm, rmax, L) in cases:
        da, db = 0.37, 1.21
        sa, sb = (0.11 if ea else 0.0), (-0.23 if eb else 0.0)
        M = 0.7
        env = (lambda r, M=M, L=L: M / (r ** (L + 1)) if r > 1e-6 else 0.0)
        v = OI.dwba_outer_integral_1d(env, ka, la, da, ea, sa, kb, lb, db, eb, sb, rm, rmax, delta_phi=np.pi / 4)
        a = OI._analytical_multipole_tail(rmax, ka, kb, da, db, la, lb, L, bound_overlap=M, eta_i=ea, eta_f=eb,
                                          sigma_i=sa, sigma_f=sb)
        tuples.append([ka, la, da, ea, sa, kb, lb, db, eb, sb, rm, rmax, L, M])
        vals.append(v)
        analytic.append(a)
    node_cases = [(1e-3, 80.0, 1.7, np.pi / 8, -0.9), (0.01, 150.0, 4.0, np.pi / 4, 1.3), (0.5, 20.0, 0.6, np.pi / 16, -2.0)]
    nodes = [OI.generate_phase_nodes(a, b, k, d, eta_total=e) for (a, b, k, d, e) in node_cases]
    np.savez_compressed(os.path.join(HERE, "tail_vectors.npz"), tuples=np.array(tuples), outer=np.array(vals),
                        analytic=np.array(analytic), node_cases=np.array(node_cases),
                        nodes=np.array(nodes, dtype=object) if False else np.concatenate(nodes),
                        node_counts=np.array([len(n) for n in nodes]))
    print(f"done {time.time() - t0:.1f}s", flush=True)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
