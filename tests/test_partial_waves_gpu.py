"""GPU checks of the batched partial-wave sums (dwba_amd.partial_waves).

* On the C1 integrals (the reference's own, 220 pairs): the device loop's
  totals equal the reference's driver-order F, G, DCS and TCS, and its
  per-l_i contributions and stopping point equal the oracle loop's.
* End to end on small synthetic scans (waves with gaps): every pair of every
  energy in one batched launch; the oracle loop runs on integrals from the
  single-pair drop-in (itself checked against the oracle) for the same pairs.
Tolerances: 1e-12 relative (FP64; the same sums in another order).
"""
import os

import numpy as np
import pytest

from oracle import partial_waves as OP

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


class _W:
    pass


def _c1_setup(cut=None):
    import torch
    from dwba_amd import partial_waves as PW
    c = np.load(os.path.join(GOLD, "c1_h1s2p_50eV.npz"))
    g = np.load(os.path.join(GOLD, "amplitudes_c1.npz"))
    I = np.array(c["I_full_split_cc5"])
    pairs = [(int(a), int(b)) for a, b in c["pairs"]]
    if cut is not None:
        for q, (li, lf) in enumerate(pairs):
            if li >= cut:
                I[q] = 0.0
    idx = {p: q for q, p in enumerate(pairs)}
    return torch, PW, c, g, I, pairs, idx


def _energy(PW, L_max_proj, k_i, k_f, chi_i_ok, chi_f_ok):
    w = _W()
    return PW.ExcitationEnergy(grid=None, V_core=None, U_i=None, U_f=None, orb_i=None, orb_f=None, k_i=k_i, k_f=k_f,
                               chi_i={l: (w if chi_i_ok(l) else None) for l in range(L_max_proj + 1)},
                               chi_f={l: (w if chi_f_ok(l) else None) for l in range(L_max_proj + 40)},
                               L_max_proj=L_max_proj)


@pytest.mark.parametrize("cut,L_proj", [(None, 20), (10, 30)])
def test_c1_loop_matches_reference_and_oracle(cut, L_proj):
    torch, PW, c, g, I, pairs, idx = _c1_setup(cut)
    chan = PW.ExcitationChannel(0, 1, 4, 1.0)
    k_i, k_f = float(g["k_i"]), float(g["k_f"])
    en = _energy(PW, L_proj, k_i, k_f, lambda l: l <= 20, lambda l: l <= 20)
    triples, has = PW.excitation_pairs([en], chan)
    rows = np.array([idx[(li, lf)] for (_, li, lf) in triples])
    dev = torch.as_tensor(I, device="cuda").contiguous()
    res, = PW.excitation_sums_from_integrals(dev, rows, triples, has, [en], chan, g["theta"])

    def integ(li, lf):
        q = idx[(li, lf)]
        return {L: float(I[q, 0, L]) for L in range(5)}, {L: float(I[q, 1, L]) for L in range(5)}
    ref = OP.excitation_loop(g["theta"], integ, lambda l: l <= 20, lambda l: l <= 20, k_i, k_f, 0, 1, 4, L_proj)
    assert res.stop_L == ref["stop_L"]
    assert (res.stop_reason is None) == (ref["stop_reason"] is None)
    sc = max(np.abs(ref["F"]).max(), np.abs(ref["G"]).max())
    assert np.abs(res.F - ref["F"]).max() <= 1e-12 * sc
    assert np.abs(res.G - ref["G"]).max() <= 1e-12 * sc
    assert np.abs(res.dcs - ref["dcs"]).max() <= 1e-12 * ref["dcs"].max()
    assert abs(res.sigma_au - ref["sigma_au"]) <= 1e-12 * ref["sigma_au"]
    assert sorted(res.partial_waves) == sorted(ref["partial_waves"])
    for k, v in ref["partial_waves"].items():
        assert abs(res.partial_waves[k] - v) <= 1e-12 * max(abs(v), 1e-300), k
    if cut is None:
        assert np.abs(res.F - g["F"]).max() <= 1e-12 * sc
        assert abs(res.sigma_au - float(g["tcs"])) <= 1e-12 * float(g["tcs"])
    else:
        assert res.stop_L == 16


def _synthetic_excitation(PW, n_e=2, seed=11):
    from dwba_amd import synthetic as S
    from dwba_amd.grid import exp_grid
    rng = np.random.default_rng(seed)
    out = []
    for e in range(n_e):
        n = int(rng.integers(1500, 2500))
        gr = exp_grid(n, 1e-5, 1e3)
        r = gr.r
        V, U = S.h_potentials(r)
        ki, kf = rng.uniform(0.8, 3.0, 2)
        Lp = 6 + e
        chi_i = {l: S.make_wave(r, ki, l, rng.uniform(0, np.pi)) for l in range(Lp + 1)}
        chi_f = {l: S.make_wave(r, kf, l, rng.uniform(0, np.pi)) for l in range(Lp + 5)}
        chi_i[2] = None                    # a failed incident solve: skipped
        chi_f[3] = None                    # a failed final solve: that l_f is left out
        out.append(PW.ExcitationEnergy(grid=gr, V_core=V, U_i=U, U_f=U, orb_i=S.Bound(S.u1s(r), l=0, energy_au=-0.5),
                                       orb_f=S.Bound(S.u2p(r), l=1, energy_au=-0.125), k_i=float(ki), k_f=float(kf),
                                       chi_i=chi_i, chi_f=chi_f, L_max_proj=Lp))
    return out


def test_excitation_scan_end_to_end():
    from dwba_amd import partial_waves as PW
    from dwba_amd.dwba_matrix_elements import radial_ME_all_L_gpu
    chan = PW.ExcitationChannel(0, 1, 4, 1.0)
    ens = _synthetic_excitation(PW)
    theta = np.linspace(0.0, np.pi, 61)
    got = PW.excitation_partial_wave_sums(ens, chan, theta, oscillatory_method="full_split")
    for en, res in zip(ens, got):
        def integ(li, lf):
            x = radial_ME_all_L_gpu(en.grid, en.V_core, en.U_i, en.orb_i, en.orb_f, en.chi_i[li], en.chi_f[lf], 4,
                                    oscillatory_method="full_split", U_f_array=en.U_f)
            return x.I_L_direct, x.I_L_exchange
        ref = OP.excitation_loop(theta, integ, lambda l: en.chi_i.get(l) is not None,
                                 lambda l: en.chi_f.get(l) is not None, en.k_i, en.k_f, 0, 1, 4, en.L_max_proj)
        assert res.skipped_li == ref["skipped"] == 1
        sc = max(np.abs(ref["F"]).max(), np.abs(ref["G"]).max())
        assert np.abs(res.F - ref["F"]).max() <= 1e-11 * sc
        assert np.abs(res.G - ref["G"]).max() <= 1e-11 * sc
        assert abs(res.sigma_au - ref["sigma_au"]) <= 1e-11 * ref["sigma_au"]
        for k, v in ref["partial_waves"].items():
            assert abs(res.partial_waves[k] - v) <= 1e-11 * max(abs(v), 1e-300), k


def test_ionization_scan_end_to_end():
    from dwba_amd import partial_waves as PW
    from dwba_amd import synthetic as S
    from dwba_amd.dwba_matrix_elements import radial_ME_all_L_gpu
    from dwba_amd.grid import exp_grid
    rng = np.random.default_rng(4)
    chan = PW.IonizationChannel(l_i=0, l_eject_max=2, L_max=4, N_equiv=1.0, convergence_threshold=0.01)
    pts = []
    for p in range(2):
        n = int(rng.integers(1500, 2500))
        gr = exp_grid(n, 1e-5, 1e3)
        r = gr.r
        V, U = S.h_potentials(r)
        ki, ke = 2.0, float(rng.uniform(0.4, 1.2))
        ks = float(np.sqrt(ki ** 2 - ke ** 2 - 0.36))
        Lp = 5
        pts.append(PW.IonizationPoint(
            grid=gr, V_core=V, U_inc=U, orb_i=S.Bound(S.slater(r, 0.6), l=0, energy_au=-0.18), k_i=ki, k_scatt=ks,
            k_eject=ke, chi_i={l: S.make_wave(r, ki, l, rng.uniform(0, np.pi)) for l in range(Lp + 1)},
            chi_eject={l: (None if l == 1 and p == 1 else S.make_wave(r, ke, l, rng.uniform(0, np.pi)))
                       for l in range(3)},
            chi_scatt={l: S.make_wave(r, ks, l, rng.uniform(0, np.pi)) for l in range(Lp + 6)},
            L_max_proj=Lp, l_f_max_scatt=Lp + 3))
    got = PW.ionization_sdcs(pts, chan, oscillatory_method="full_split")
    for pt, res in zip(pts, got):
        def integ(li, le, lf):
            x = radial_ME_all_L_gpu(pt.grid, pt.V_core, pt.U_inc, pt.orb_i, pt.chi_eject[le], pt.chi_i[li],
                                    pt.chi_scatt[lf], 4, oscillatory_method="full_split", U_f_array=pt.V_core)
            return x.I_L_direct, x.I_L_exchange
        ref = OP.ionization_loop(integ, lambda l: pt.chi_i.get(l) is not None,
                                 lambda l: pt.chi_eject.get(l) is not None, lambda l: pt.chi_scatt.get(l) is not None,
                                 pt.k_i, pt.k_scatt, pt.k_eject, 0, 2, 4, pt.L_max_proj, pt.l_f_max_scatt)
        assert res.last_l_i == ref["last_l_i"]
        assert abs(res.sigma_au - ref["sigma_au"]) <= 1e-11 * ref["sigma_au"]
        for k, v in ref["partial_L"].items():
            assert abs(res.partial_L[k] - v) <= 1e-11 * max(abs(v), 1e-300), k
        for (le, lf, mi, mf), (f, g) in ref["coherent"].items():
            c = res.coherent[(le, lf)][res.combos.index((mi, mf))]
            sc = max(abs(f), abs(g))
            assert abs(c[0] - f) <= 1e-11 * sc and abs(c[1] - g) <= 1e-11 * sc


def test_c1_record_in_reference_format(tmp_path):
    """From the device sums to a stored record: top-up on the per-l_i sums,
    sigma in both units, DCS on theta in degrees, merged into results/."""
    import json
    from dwba_amd import results as R
    torch, PW, c, g, I, pairs, idx = _c1_setup()
    chan = PW.ExcitationChannel(0, 1, 4, 1.0)
    en = _energy(PW, 20, float(g["k_i"]), float(g["k_f"]), lambda l: l <= 20, lambda l: l <= 20)
    triples, has = PW.excitation_pairs([en], chan)
    rows = np.array([idx[(li, lf)] for (_, li, lf) in triples])
    res, = PW.excitation_sums_from_integrals(torch.as_tensor(I, device="cuda").contiguous(), rows, triples, has,
                                             [en], chan, g["theta"])
    entry = R.excitation_result_entry(50.0, res, g["theta"], 0, 1, 4, 20, run_meta={"model": "static"})
    want_sig, want_pw, top = R.excitation_topup(res.partial_waves, R.sigma_au_to_cm2(float(g["tcs"])), 0, 1)
    assert abs(entry["sigma_cm2"] - want_sig) <= 1e-12 * want_sig
    assert abs(entry["sigma_au"] - R.sigma_cm2_to_au(entry["sigma_cm2"])) <= 1e-12 * entry["sigma_au"]
    assert entry["meta"]["runtime"]["n_projectile_partial_waves_summed"] == 21
    assert np.allclose(entry["theta_deg"], np.degrees(g["theta"]), rtol=0, atol=1e-12)
    p = R.save_results("results_H2p_exc.json", {"H_1s-2p": [entry]}, root=tmp_path)
    back = json.loads(p.read_text())["H_1s-2p"][0]
    assert back["partial_waves"] == entry["partial_waves"] and back["dcs_au_raw"] == entry["dcs_au_raw"]
