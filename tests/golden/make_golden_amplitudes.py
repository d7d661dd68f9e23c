"""Write the golden vectors of the reference's amplitude assembly (run once, in
the container that has the reference checked out; the .npz is committed).

The composite amplitudes -- the phase conventions, the mu_f selection of the
direct and exchange terms, the spherical-harmonic conventions -- are pinned
by nothing but the reference itself, so this script imports its
dwba_coupling and sigma_total modules and records, on the C1 inputs
(H 1s -> 2p at 50 eV, the 220 pairs l_i + l_f odd, L = 0..4, I_L from
c1_h1s2p_50eV.npz, which the reference's CPU path wrote):

  theta (181 points on [0, pi]), k_i, k_f;
  f, g of calculate_amplitude_contribution for (Mi, Mf) = (0, -1), (0, 0), (0, 1)
     summed over all 220 pairs (the driver's accumulation, driver.py:1041-1052),
     and per pair for 20 pairs on 19 angles;
  dcs_dwba of the sums (N_equiv = 1) summed over Mf, and integrate_dcs_over_angles;
  calculate_ionization_coefficients for 40 seeded (l_i, l_f, l_ej, L_i, M_i, M_f)
     with seeded I_L (L = 0..6) and k's;
  wigner_3j_num, wigner_6j_num, clebsch_gordan, racah_W on 300 seeded arguments each.

Usage:  python tests/golden/make_golden_amplitudes.py [/root/reference]
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main(ref_root: str) -> None:
    sys.path.insert(0, ref_root)
    os.environ.setdefault("DWBA_LOG_LEVEL", "ERROR")
    import warnings
    warnings.filterwarnings("ignore")
    import dwba_coupling as DC
    import sigma_total as ST

    c1 = np.load(os.path.join(HERE, "c1_h1s2p_50eV.npz"))
    k_i = float(c1["meta_i"][0][0])
    k_f = float(c1["meta_f"][0][0])
    pairs = [tuple(int(x) for x in p) for p in c1["pairs"]]
    I = c1["I_full_split_cc5"]
    L_max = int(c1["L_max"])
    theta = np.linspace(0.0, np.pi, 181)
    theta_s = np.linspace(0.0, np.pi, 19)
    Mfs = (-1, 0, 1)
    F = np.zeros((3, theta.size), complex)
    G = np.zeros((3, theta.size), complex)
    sub = list(range(0, len(pairs), 11))
    f_sub = np.zeros((len(sub), 3, theta_s.size), complex)
    g_sub = np.zeros((len(sub), 3, theta_s.size), complex)
    for q, (li, lf) in enumerate(pairs):
        d = {L: float(I[q, 0, L]) for L in range(L_max + 1)}
        e = {L: float(I[q, 1, L]) for L in range(L_max + 1)}
        for c, Mf in enumerate(Mfs):
            a = DC.calculate_amplitude_contribution(theta, d, e, li, lf, k_i, k_f, 0, 1, 0, Mf)
            F[c] += a.f_theta
            G[c] += a.g_theta
            if q in sub:
                b = DC.calculate_amplitude_contribution(theta_s, d, e, li, lf, k_i, k_f, 0, 1, 0, Mf)
                f_sub[sub.index(q), c] = b.f_theta
                g_sub[sub.index(q), c] = b.g_theta
    dcs = np.zeros(theta.size)
    for c in range(3):
        dcs += ST.dcs_dwba(theta, F[c], G[c], k_i, k_f, 0, 1)
    tcs = ST.integrate_dcs_over_angles(theta, dcs)

    rng = np.random.default_rng(77)
    ion_args, ion_I, ion_out = [], [], []
    while len(ion_args) < 40:
        li, lf, le = (int(x) for x in rng.integers(0, 9, 3))
        Li = int(rng.integers(0, 3))
        Mi = int(rng.integers(-Li, Li + 1))
        Mf = int(rng.integers(-le, le + 1))
        ki, kf, ke = rng.uniform(0.3, 3.0, 3)
        Id = rng.standard_normal(7)
        Ie = rng.standard_normal(7)
        Id[rng.random(7) < 0.2] = 1e-22       # below the reference's 1e-20 cut
        res = DC.calculate_ionization_coefficients({L: float(Id[L]) for L in range(7)},
                                                   {L: float(Ie[L]) for L in range(7)},
                                                   li, lf, le, ki, kf, ke, Li, Mi, Mf)
        ion_args.append([li, lf, le, Li, Mi, Mf, ki, kf, ke])
        ion_I.append([Id, Ie])
        ion_out.append([res.f_coeff, res.g_coeff])

    w3, w6, cg, rw = [], [], [], []
    for _ in range(300):
        j1, j2 = (int(x) for x in rng.integers(0, 12, 2))
        j3 = int(rng.integers(abs(j1 - j2), j1 + j2 + 1))
        m1 = int(rng.integers(-j1, j1 + 1))
        m2 = int(rng.integers(-j2, j2 + 1))
        m3 = -m1 - m2
        w3.append([j1, j2, j3, m1, m2, m3, DC.wigner_3j_num(j1, j2, j3, m1, m2, m3)])
        cg.append([j1, j2, j3, m1, m2, m1 + m2, DC.clebsch_gordan(j1, j2, j3, m1, m2, m1 + m2)])
        a = [int(x) for x in rng.integers(0, 9, 6)]
        w6.append(a + [DC.wigner_6j_num(*a)])
        rw.append(a + [DC.racah_W(*a)])

    np.savez_compressed(
        os.path.join(HERE, "amplitudes_c1.npz"), theta=theta, theta_sub=theta_s, k_i=k_i, k_f=k_f,
        Mf=np.array(Mfs), F=F, G=G, sub_pairs=np.array(sub), f_sub=f_sub, g_sub=g_sub, dcs=dcs, tcs=tcs,
        ion_args=np.array(ion_args), ion_I=np.array(ion_I), ion_out=np.array(ion_out),
        w3j=np.array(w3), w6j=np.array(w6), cg=np.array(cg), racah=np.array(rw))
    print("tcs", tcs, "max|F|", np.abs(F).max(), "max|G|", np.abs(G).max())


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
