"""Golden vectors of the reference's own CPU path (tests/golden/make_golden.py):
H 1s -> 2p at 50 eV, Z = 1 Coulomb SAE core, make_r_grid(1e-4, 200, 3000),
continuum waves l <= 20 from the reference's Numerov solver, the 220 pairs with
l_i + l_f odd, L = 0..4 (full_split, CC_nodes = 5), plus full_split CC 7,
advanced and legacy on 20 pairs; outer tails (dwba_outer_integral_1d) and the
analytical tail on 20 parameter tuples; Coulomb phase nodes.

CPU tests pin the oracle to these vectors; GPU tests hold the drop-in and the
batched engine to them.  Bar (FP64): |I - I_ref| <= 1e-10 |I_ref| + 1e-13 M
+ 10 x the Levin-branch noise estimate of the outer tail (M: cancellation-free
magnitude of the same integral)."""
import os

import numpy as np
import pytest

from oracle import quadrature as Q, radial as R, semisep, tail as T
from dwba_amd.synthetic import Bound, Wave

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
C1 = np.load(os.path.join(HERE, "c1_h1s2p_50eV.npz"))
TV = np.load(os.path.join(HERE, "tail_vectors.npz"))
LMAX = int(C1["L_max"])


class _Grid:
    def __init__(self, d):
        self.r, self.w_simpson, self.w_trapz = d["r"], d["w_simpson"], d["w_trapz"]


def _wave(meta, chi):
    k, l, ps, im, eta, sig = meta
    return Wave(l=int(l), k_au=float(k), chi_of_r=chi, phase_shift=float(ps), idx_match=int(im), eta=float(eta),
                sigma_l=float(sig))


GRID = _Grid(C1)
BI, BF = Bound(C1["u_i"], l=0, energy_au=float(C1["e_i"])), Bound(C1["u_f"], l=1, energy_au=float(C1["e_f"]))
WI = [_wave(C1["meta_i"][l], C1["chi_i"][l]) for l in range(C1["chi_i"].shape[0])]
WF = [_wave(C1["meta_f"][l], C1["chi_f"][l]) for l in range(C1["chi_f"].shape[0])]
PAIRS = [tuple(int(x) for x in p) for p in C1["pairs"]]
VARIANTS = [(str(m), int(c)) for m, c in zip(C1["variant_methods"], C1["variant_cc"])]


def _oracle(q, method="full_split", cc=5):
    li, lf = PAIRS[q]
    return R.radial_ME_all_L_reduced(GRID.r, GRID.w_simpson, GRID.w_trapz, C1["V_core"], C1["U_i"], BI, BF, WI[li],
                                     WF[lf], LMAX, oscillatory_method=method, CC_nodes=cc, U_f_array=C1["U_f"],
                                     return_parts=True)


def _tol(parts, ref_fam, L, fam, rtol=1e-10, atol_rel=1e-13):
    a, b = (parts["alpha_dir"], parts["beta_dir"]) if fam == 0 else (parts["alpha_ex"], parts["beta_ex"])
    mag = semisep.il_all_L(GRID.r, np.abs(a), np.abs(b), LMAX)[L]
    if fam == 0:
        mag += abs(parts["tails"][L])
    noise = 10.0 * parts["tail_noise"][L] if fam == 0 else 0.0
    return rtol * abs(ref_fam) + atol_rel * mag + noise


def _assert_matches(got_dir, got_ex, gold, parts, label, **kw):
    for L in range(LMAX + 1):
        for fam, got in ((0, got_dir[L]), (1, got_ex[L])):
            ref = gold[fam, L]
            assert abs(got - ref) <= _tol(parts, ref, L, fam, **kw), (label, fam, L, got, ref)


# ------------------------------------------------------------------ CPU: the oracle against the reference

def test_oracle_reduced_matches_reference_all_pairs():
    gold = C1["I_full_split_cc5"]
    for q in range(len(PAIRS)):
        res, parts = _oracle(q)
        _assert_matches(res.I_L_direct, res.I_L_exchange, gold[q], parts, PAIRS[q])


@pytest.mark.parametrize("v", range(3))
def test_oracle_reduced_matches_reference_variants(v):
    method, cc = VARIANTS[v]
    gold = C1["I_variants"][v]
    for s_i, q in enumerate(C1["variant_pairs"]):
        res, parts = _oracle(int(q), method, cc)
        _assert_matches(res.I_L_direct, res.I_L_exchange, gold[s_i], parts, (method, cc, PAIRS[int(q)]))


@pytest.mark.parametrize("q", [0, 37, 110, 219])
def test_oracle_faithful_matches_reference(q):
    li, lf = PAIRS[q]
    res = R.radial_ME_all_L_faithful(GRID.r, GRID.w_simpson, GRID.w_trapz, C1["V_core"], C1["U_i"], BI, BF, WI[li],
                                     WF[lf], LMAX, oscillatory_method="full_split", CC_nodes=5, U_f_array=C1["U_f"])
    _, parts = _oracle(q)
    _assert_matches(res.I_L_direct, res.I_L_exchange, C1["I_full_split_cc5"][q], parts, PAIRS[q])


def test_oracle_outer_tails_match_reference():
    for t, tup in enumerate(TV["tuples"]):
        ka, la, da, ea, sa, kb, lb, db, eb, sb, rm, rmax, L, M = tup
        la, lb, L = int(la), int(lb), int(L)
        env = (lambda r, M=M, L=L: M / (r ** (L + 1)) if r > 1e-6 else 0.0)
        got = T.dwba_outer_integral(env, ka, la, da, ea, sa, kb, lb, db, eb, sb, rm, rmax, delta_phi=np.pi / 4)
        noise = T.outer_integral_noise(env, ka, la, da, ea, sa, kb, lb, db, eb, sb, rm, rmax, delta_phi=np.pi / 4)
        ref = TV["outer"][t]
        assert abs(got - ref) <= 1e-10 * abs(ref) + 10.0 * noise + 1e-300, (t, got, ref, noise)
        an = T.analytical_multipole_tail(rmax, ka, kb, da, db, la, lb, L, bound_overlap=M, eta_i=ea, eta_f=eb,
                                         sigma_i=sa, sigma_f=sb)
        assert abs(an - TV["analytic"][t]) <= 1e-12 * abs(TV["analytic"][t]) + 1e-300, (t, an, TV["analytic"][t])


def test_oracle_phase_nodes_match_reference():
    off = 0
    for c, (a, b, k, dphi, eta) in enumerate(TV["node_cases"]):
        n = int(TV["node_counts"][c])
        ref = TV["nodes"][off:off + n]
        off += n
        got = Q.generate_phase_nodes(a, b, k, dphi, eta_total=eta)
        assert len(got) == n
        np.testing.assert_allclose(got, ref, rtol=1e-13, atol=0)


# ------------------------------------------------------------------ GPU: the product against the reference

@pytest.mark.gpu
@pytest.mark.parametrize("contraction", ["semiseparable", "dense"])
def test_dropin_matches_reference_all_pairs(contraction):
    from dwba_amd import dwba_matrix_elements as ME
    print(f"\n[golden] drop-in, 220 pairs, contraction={contraction}", flush=True)
    prev = ME.get_contraction()
    ME.set_contraction(contraction)
    try:
        cache = ME.GPUCache.from_grid(GRID)
        for q, (li, lf) in enumerate(PAIRS):
            res = ME.radial_ME_all_L_gpu(GRID, C1["V_core"], C1["U_i"], BI, BF, WI[li], WF[lf], LMAX,
                                         oscillatory_method="full_split", CC_nodes=5, gpu_cache=cache,
                                         U_f_array=C1["U_f"])
            _, parts = _oracle(q)
            _assert_matches(res.I_L_direct, res.I_L_exchange, C1["I_full_split_cc5"][q], parts, PAIRS[q])
    finally:
        ME.set_contraction(prev)


@pytest.mark.gpu
@pytest.mark.parametrize("v", range(3))
def test_dropin_matches_reference_variants(v):
    from dwba_amd import dwba_matrix_elements as ME
    method, cc = VARIANTS[v]
    print(f"\n[golden] drop-in variant {method} CC{cc}", flush=True)
    for s_i, q in enumerate(C1["variant_pairs"]):
        li, lf = PAIRS[int(q)]
        res = ME.radial_ME_all_L(GRID, C1["V_core"], C1["U_i"], BI, BF, WI[li], WF[lf], LMAX,
                                 oscillatory_method=method, CC_nodes=cc, U_f_array=C1["U_f"])
        _, parts = _oracle(int(q), method, cc)
        _assert_matches(res.I_L_direct, res.I_L_exchange, C1["I_variants"][v][s_i], parts, (method, cc, li, lf))


@pytest.mark.gpu
@pytest.mark.parametrize("precision,contraction,rtol,atol_rel", [("fp64", None, 1e-10, 1e-13),
                                                                 ("fp32", "block", 0.0, 1e-6),
                                                                 ("fp32", "dense", 0.0, 1e-6),
                                                                 ("bf16x3", "block", 0.0, 3e-6),
                                                                 ("bf16x3", "dense", 0.0, 1e-6)])
def test_batch_matches_reference(precision, contraction, rtol, atol_rel):
    from dwba_amd.batch import ScanBatch, EnergyInput
    print(f"\n[golden] batch precision={precision} contraction={contraction}", flush=True)
    en = EnergyInput(grid=GRID, V_core=C1["V_core"], U_i=C1["U_i"], bound_i=BI, bound_f=BF,
                     pairs=[(WI[li], WF[lf]) for (li, lf) in PAIRS], U_f=C1["U_f"])
    b = ScanBatch([en], LMAX, precision=precision, oscillatory_method="full_split", CC_nodes=5,
                  contraction=contraction)
    b.run()
    res = b.results()
    for q in range(len(PAIRS)):
        _, parts = _oracle(q)
        _assert_matches(res[q, 0], res[q, 1], C1["I_full_split_cc5"][q], parts, (precision, PAIRS[q]), rtol=rtol,
                        atol_rel=atol_rel)
