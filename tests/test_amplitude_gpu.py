"""GPU checks of amplitude assembly (csrc/amplitude.hip through the C-ABI).

Against the reference's own outputs (tests/golden/amplitudes_c1.npz, from
dwba_coupling.py / sigma_total.py on the C1 integrals): the driver-summed
F, G over 220 pairs and 3 projection pairs on 181 angles, per-pair f, g, the
DCS and TCS, and 40 ionization coefficient sets.  Against the oracle: the
batched assembly fed straight from the scan's device output, per-slot sums
and their linearity.  Tolerances: 1e-12 of the largest magnitude (FP64, the
same sums in another order).
"""
import math
import os

import numpy as np
import pytest

from oracle import amplitude as OA

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def gold():
    return np.load(os.path.join(GOLD, "amplitudes_c1.npz"))


@pytest.fixture(scope="module")
def c1():
    return np.load(os.path.join(GOLD, "c1_h1s2p_50eV.npz"))


def _rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300))


def test_ylm_table_matches_scipy():
    import torch
    from dwba_amd.amplitudes import AmplitudeAssembler
    theta = np.linspace(0.0, np.pi, 73)
    asm = AmplitudeAssembler(theta, 3, 3, 4)                 # m_max = 6
    y = asm.ylm(45).view(46, 13, theta.size).cpu().numpy()
    for l in range(46):
        for m in range(-6, 7):
            want = OA.ylm(l, m, theta).real if abs(m) <= l else 0.0 * theta
            assert np.abs(y[l, m + 6] - want).max() <= 1e-13 * max(1.0, np.abs(want).max()), (l, m)
    torch.cuda.synchronize()


def test_driver_sums_match_reference(gold, c1):
    import torch
    from dwba_amd.amplitudes import AmplitudeAssembler
    theta = gold["theta"]
    I = torch.as_tensor(c1["I_full_split_cc5"], device="cuda").contiguous()
    L_max = I.shape[2] - 1
    combos = [(0, int(m)) for m in gold["Mf"]]
    asm = AmplitudeAssembler(theta, 0, 1, L_max, combos=combos)
    pairs = [(q, int(li), int(lf), 0) for q, (li, lf) in enumerate(c1["pairs"])]
    res = asm.excitation(I, pairs, [float(gold["k_i"])], [float(gold["k_f"])])
    F, G = res.F[0].cpu().numpy(), res.G[0].cpu().numpy()
    scale = max(np.abs(gold["F"]).max(), np.abs(gold["G"]).max())
    assert np.abs(F - gold["F"]).max() <= 1e-12 * scale
    assert np.abs(G - gold["G"]).max() <= 1e-12 * scale
    assert _rel(res.dcs[0].cpu().numpy(), gold["dcs"]) <= 1e-12
    assert abs(float(res.tcs[0]) - float(gold["tcs"])) <= 1e-12 * float(gold["tcs"])


def test_dropin_per_pair(gold, c1):
    from dwba_amd.amplitudes import calculate_amplitude_contribution
    theta = gold["theta_sub"]
    I = c1["I_full_split_cc5"]
    for s, q in enumerate(gold["sub_pairs"]):
        li, lf = (int(x) for x in c1["pairs"][q])
        d = {L: float(I[q, 0, L]) for L in range(I.shape[2])}
        e = {L: float(I[q, 1, L]) for L in range(I.shape[2])}
        for c, Mf in enumerate(gold["Mf"]):
            a = calculate_amplitude_contribution(theta, d, e, li, lf, float(gold["k_i"]), float(gold["k_f"]),
                                                 0, 1, 0, int(Mf))
            sc = max(np.abs(gold["f_sub"][s, c]).max(), np.abs(gold["g_sub"][s, c]).max(), 1e-300)
            assert np.abs(a.f_theta - gold["f_sub"][s, c]).max() <= 1e-12 * sc, (q, Mf)
            assert np.abs(a.g_theta - gold["g_sub"][s, c]).max() <= 1e-12 * sc, (q, Mf)


def test_dcs_dropins(gold):
    from dwba_amd.amplitudes import dcs_dwba, integrate_dcs_over_angles
    theta = gold["theta"]
    d = sum(dcs_dwba(theta, gold["F"][c], gold["G"][c], float(gold["k_i"]), float(gold["k_f"]), 0, 1)
            for c in range(3))
    assert _rel(d, gold["dcs"]) <= 1e-13
    assert abs(integrate_dcs_over_angles(theta, gold["dcs"]) - float(gold["tcs"])) <= 1e-13 * float(gold["tcs"])
    z = dcs_dwba(theta, gold["F"][0], gold["G"][0], 0.0, 1.0, 0, 1)
    assert not np.any(z)


def test_ionization_coefficients(gold):
    import torch
    from dwba_amd.amplitudes import IonizationAssembler, calculate_ionization_coefficients
    for args, Iv, want in zip(gold["ion_args"], gold["ion_I"], gold["ion_out"]):
        li, lf, le, Li, Mi, Mf = (int(x) for x in args[:6])
        ki, kf, ke = (float(x) for x in args[6:])
        d = {L: float(Iv[0][L]) for L in range(Iv.shape[1])}
        e = {L: float(Iv[1][L]) for L in range(Iv.shape[1])}
        c = calculate_ionization_coefficients(d, e, li, lf, le, ki, kf, ke, Li, Mi, Mf)
        sc = max(abs(want[0]), abs(want[1]), 1e-300)
        assert abs(c.f_coeff - want[0]) <= 1e-12 * sc and abs(c.g_coeff - want[1]) <= 1e-12 * sc
    # batched: all 40 in one launch, grouped by L_i
    for Li in sorted({int(a[3]) for a in gold["ion_args"]}):
        sel = [i for i, a in enumerate(gold["ion_args"]) if int(a[3]) == Li]
        asm = IonizationAssembler(Li, 8, gold["ion_I"].shape[2] - 1)
        I = torch.as_tensor(np.ascontiguousarray(gold["ion_I"][sel]), device="cuda")
        A = gold["ion_args"][sel]
        out = asm.coefficients(I, [(q, int(a[0]), int(a[1]), int(a[2])) for q, a in enumerate(A)],
                               A[:, 6], A[:, 7], A[:, 8]).cpu().numpy()
        for q, i in enumerate(sel):
            k = asm.combos.index((int(A[q][4]), int(A[q][5])))
            want = gold["ion_out"][i]
            sc = max(abs(want[0]), abs(want[1]), 1e-300)
            assert np.abs(out[q, k] - want).max() <= 1e-12 * sc
            for kk, (mi, mf) in enumerate(asm.combos):      # |M_f| > l_ej: no such state
                if abs(mf) > int(A[q][2]):
                    assert out[q, kk, 0] == 0 and out[q, kk, 1] == 0


def test_ionization_coherent_sums_are_slot_sums(gold):
    import torch
    from dwba_amd.amplitudes import IonizationAssembler
    sel = [i for i, a in enumerate(gold["ion_args"]) if int(a[3]) == 1]
    A = gold["ion_args"][sel]
    asm = IonizationAssembler(1, 8, gold["ion_I"].shape[2] - 1)
    I = torch.as_tensor(np.ascontiguousarray(gold["ion_I"][sel]), device="cuda")
    pairs = [(q, int(a[0]), int(a[1]), int(a[2])) for q, a in enumerate(A)]
    each = asm.coefficients(I, pairs, A[:, 6], A[:, 7], A[:, 8]).cpu().numpy()
    slot = np.arange(len(sel)) % 3
    sums = asm.coherent_sums(I, pairs, A[:, 6], A[:, 7], A[:, 8], slot).cpu().numpy()
    for s in range(3):
        want = each[slot == s].sum(axis=0)
        assert np.abs(sums[s] - want).max() <= 1e-13 * max(np.abs(want).max(), 1e-300)


def test_scan_output_to_cross_sections():
    """Integrals left on the device by the batched scan feed the assembly
    directly; per-energy and per-(energy, l_i) slots against the oracle on
    the same integrals."""
    import torch
    from dwba_amd import synthetic as S
    from dwba_amd.amplitudes import AmplitudeAssembler
    from dwba_amd.batch import ScanBatch
    scan = S.excitation_scan(n_energies=3, n_pairs=24, n_range=(1500, 2500), l_max=12, k_range=(0.8, 3.0), seed=5)
    L_max = 6
    b = ScanBatch(scan, L_max, precision="fp64", oscillatory_method="full_split")
    b.run()
    out = b.out
    I_h = out.cpu().numpy()
    theta = np.linspace(0.0, np.pi, 91)
    asm = AmplitudeAssembler(theta, 0, 1, L_max)
    k_i = [en.pairs[0][0].k_au for en in scan]
    k_f = [en.pairs[0][1].k_au for en in scan]
    # pair k of the scan is stored at row out_index[k] of the device output
    rows = [(int(b.out_index[k]), pr[3].l, pr[4].l, pr[0]) for k, pr in enumerate(b.pairs)]
    res = asm.excitation(out, rows, k_i, k_f)
    # the oracle on the same integrals, energy by energy
    for e in range(len(scan)):
        F = np.zeros((3, theta.size), complex)
        G = np.zeros_like(F)
        for (k, li, lf, ei) in rows:
            if ei != e:
                continue
            d = {L: float(I_h[k, 0, L]) for L in range(L_max + 1)}
            x = {L: float(I_h[k, 1, L]) for L in range(L_max + 1)}
            for c, (mi, mf) in enumerate(asm.combos):
                f, g = OA.amplitude_contribution(theta, d, x, li, lf, k_i[e], k_f[e], 0, 1, mi, mf)
                F[c] += f
                G[c] += g
        sc = max(np.abs(F).max(), np.abs(G).max())
        assert np.abs(res.F[e].cpu().numpy() - F).max() <= 1e-12 * sc
        assert np.abs(res.G[e].cpu().numpy() - G).max() <= 1e-12 * sc
        dcs = sum(OA.dcs_dwba(theta, F[c], G[c], k_i[e], k_f[e], 0, 1) for c in range(3))
        assert _rel(res.dcs[e].cpu().numpy(), dcs) <= 1e-12
        assert abs(float(res.tcs[e]) - OA.integrate_dcs(theta, dcs)) <= 1e-12 * OA.integrate_dcs(theta, dcs)
    # per-(energy, l_i) slots: amplitudes are linear, so the slots of an energy sum to its total
    li_max = max(r[1] for r in rows)
    rows2 = [(k, li, lf, e * (li_max + 1) + li) for (k, li, lf, e) in rows]
    ns = len(scan) * (li_max + 1)
    k_i2 = np.repeat(k_i, li_max + 1)
    k_f2 = np.repeat(k_f, li_max + 1)
    res2 = asm.excitation(out, rows2, k_i2, k_f2, n_slots=ns)
    F2 = res2.F.view(len(scan), li_max + 1, 3, -1).sum(dim=1)
    assert torch.allclose(F2, res.F, rtol=0, atol=1e-12 * float(res.F.abs().max()))
    empty = [s for s in range(ns) if s not in {r[3] for r in rows2}]
    if empty:
        assert float(res2.F[empty].abs().max()) == 0.0 and float(res2.tcs[empty].abs().max()) == 0.0


def test_bad_arguments_raise():
    import torch
    from dwba_amd.amplitudes import AmplitudeAssembler
    asm = AmplitudeAssembler(np.linspace(0, np.pi, 5), 0, 1, 4)
    I = torch.zeros((2, 2, 5), dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError):
        asm.excitation(I, [(2, 0, 1, 0)], [1.0], [1.0])          # row outside
    with pytest.raises(ValueError):
        asm.excitation(torch.zeros((2, 2, 4), dtype=torch.float64, device="cuda"), [(0, 0, 1, 0)], [1.0], [1.0])
    with pytest.raises(ValueError):
        AmplitudeAssembler(np.linspace(0, np.pi, 5), 0, 1, 64)
    with pytest.raises(ValueError):
        AmplitudeAssembler(np.linspace(0, np.pi, 5), 4, 4, 4)    # 81 projection pairs
