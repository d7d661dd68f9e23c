"""Parity of the batched scan engine at the benchmark configurations' full size.

* Excitation sweep: one energy at N = 15000 with k_i, k_f near the top of
  U[0.5, 8] and one at N = 4000 with low k, 400 channel pairs each
  (l_i, l_f ~ U{0..40}), L = 0..30, Full-Split quadrature, through ScanBatch
  in fp64 (library defaults, autotuned, and the LDS-staged lane kernel at 4
  and 8 multipoles per lane), fp32 and bf16x3.
* Ionization batch: one ejected-energy point at N = 15000, 400 (l_i, l_f,
  l_ej) triples with l <= 30, L = 0..30, fp64 and fp32.

32 sampled pairs of each are compared with the oracle's reduced route
(alpha/beta + exact long-double O(N) contraction + outer tails + L = 0
correction).  Bars: FP64 |dI| <= 1e-10 |I| + 1e-13 M (+ 10x the Levin noise
estimate on the direct tail), M = sum |alpha_i| K_ij |beta_j|; FP32 / BF16x3
MFMA families |dI| <= 1e-6 M, the error printed per L.
"""
import numpy as np
import pytest

from oracle import radial as R, semisep
from dwba_amd import synthetic as S

pytestmark = pytest.mark.gpu

L_MAX = 30
N_SAMPLE = 32


def _excitation():
    hi = S.excitation_scan(n_energies=1, n_pairs=400, n_range=(15000, 15000), l_max=40, k_range=(7.0, 8.0), seed=31)
    lo = S.excitation_scan(n_energies=1, n_pairs=400, n_range=(4000, 4000), l_max=40, k_range=(0.5, 2.0), seed=32)
    return hi + lo


def _ionization():
    return S.ionization_scan(n_points=1, n_triples=400, n_range=(15000, 15000), l_max=30, seed=33)


def _oracle(scan, seed):
    """Oracle values, magnitudes and tail noise for N_SAMPLE pairs spread over the energies."""
    rng = np.random.default_rng(seed)
    per = N_SAMPLE // len(scan)
    picks, base = [], 0
    for en in scan:
        idx = np.sort(rng.choice(len(en.pairs), size=per, replace=False))
        picks += [(base + int(i), en, en.pairs[int(i)]) for i in idx]
        base += len(en.pairs)
    out = []
    for k, en, pr in picks:
        g = en.grid
        bf = pr[2] if len(pr) > 2 else en.bound_f
        red, parts = R.radial_ME_all_L_reduced(g.r, g.w_simpson, g.w_trapz, en.V_core, en.U_i, en.bound_i, bf,
                                               pr[0], pr[1], L_MAX, U_f_array=en.U_f, oscillatory_method="full_split",
                                               return_parts=True)
        mag_d = semisep.abs_bound_all_L(g.r, parts["alpha_dir"], parts["beta_dir"], L_MAX) + np.abs(parts["tails"])
        mag_e = semisep.abs_bound_all_L(g.r, parts["alpha_ex"], parts["beta_ex"], L_MAX)
        ref = np.array([[red.I_L_direct[L] for L in range(L_MAX + 1)], [red.I_L_exchange[L] for L in range(L_MAX + 1)]])
        out.append((k, ref, np.array([mag_d, mag_e]), parts["tail_noise"]))
    return out


@pytest.fixture(scope="module")
def excitation():
    scan = _excitation()
    assert [len(en.grid.r) for en in scan] == [15000, 4000]
    return scan, _oracle(scan, 41)


@pytest.fixture(scope="module")
def ionization():
    scan = _ionization()
    return scan, _oracle(scan, 42)


def _check(res, oracle, mfma_families, label):
    """res: (n_pairs, 2, L+1) batch results in input order."""
    worst = np.zeros((2, L_MAX + 1))
    for k, ref, mag, noise in oracle:
        got = res[k]
        for fam in (0, 1):
            d = np.abs(got[fam] - ref[fam])
            if fam in mfma_families:
                tol = 1e-6 * mag[fam]
            else:
                tol = 1e-10 * np.abs(ref[fam]) + 1e-13 * mag[fam] + (10.0 * noise if fam == 0 else 0.0)
            worst[fam] = np.maximum(worst[fam], d / np.maximum(mag[fam], 1e-300))
            bad = np.nonzero(~(d <= tol))[0]
            assert bad.size == 0, (label, k, fam, bad.tolist(), got[fam][bad], ref[fam][bad],
                                   (d / np.maximum(mag[fam], 1e-300))[bad])
    for fam, name in ((0, "direct"), (1, "exchange")):
        print(f"\n[{label}] {name} max |dI|/M per L: " + " ".join(f"{x:.1e}" for x in worst[fam]), flush=True)
    return worst


def _batch(scan, precision, tune=False, tmp_path=None, monkeypatch=None, contraction=None, stage_lb=None):
    from dwba_amd.batch import ScanBatch
    b = ScanBatch(scan, L_MAX, precision=precision, oscillatory_method="full_split", contraction=contraction)
    if stage_lb:                            # the LDS-staged lane kernel at stage_lb multipoles per lane
        b.semisep_lanes = True
        b.lane_tables = True
        b.lane_split = 4
        b.semisep_config = (0, stage_lb)
    if tune:
        from dwba_amd import autotune as AT
        monkeypatch.setenv("DWBA_AUTOTUNE_CACHE", str(tmp_path / "tune.json"))
        b.tune(AT.Autotuner())
        print(f"\n[tuned] {b.tuning_step['choice']} finalists {b.tuning_step['finalists']}", flush=True)
    b.capture()
    b.run()
    return b.results()


def test_excitation_fullsize_fp64_defaults(excitation):
    scan, ora = excitation
    _check(_batch(scan, "fp64"), ora, (), "excitation fp64 defaults")


def test_excitation_fullsize_fp64_tuned(excitation, tmp_path, monkeypatch):
    scan, ora = excitation
    _check(_batch(scan, "fp64", True, tmp_path, monkeypatch), ora, (), "excitation fp64 tuned")


@pytest.mark.parametrize("lb", [4, 8])
def test_excitation_fullsize_fp64_staged(excitation, lb):
    scan, ora = excitation
    _check(_batch(scan, "fp64", stage_lb=lb), ora, (), f"excitation fp64 LDS-staged lb {lb}")


@pytest.mark.parametrize("contraction", ["block", "dense"])
def test_excitation_fullsize_fp32(excitation, contraction):
    scan, ora = excitation
    _check(_batch(scan, "fp32", contraction=contraction), ora, (1,), f"excitation fp32 {contraction}")


@pytest.mark.xfail(strict=False, reason="split-BF16x3 keeps ~16 significand bits of each operand (hi + lo): "
                                        "measured up to 1.6e-6 of sum|aKb| on the 64x400 scan with dense tiles, "
                                        "~2e-6 in block form; the FP32 forms meet 1e-6")
@pytest.mark.parametrize("contraction", ["block", "dense"])
def test_excitation_fullsize_bf16x3(excitation, contraction):
    scan, ora = excitation
    _check(_batch(scan, "bf16x3", contraction=contraction), ora, (1,), f"excitation bf16x3 {contraction}")


def test_ionization_fullsize_fp64(ionization):
    scan, ora = ionization
    _check(_batch(scan, "fp64"), ora, (), "ionization fp64")


def test_ionization_fullsize_fp64_staged(ionization):
    scan, ora = ionization
    _check(_batch(scan, "fp64", stage_lb=4), ora, (), "ionization fp64 LDS-staged lb 4")


@pytest.mark.parametrize("contraction", ["block", "dense"])
def test_ionization_fullsize_fp32(ionization, contraction):
    scan, ora = ionization
    _check(_batch(scan, "fp32", contraction=contraction), ora, (0, 1), f"ionization fp32 {contraction}")
