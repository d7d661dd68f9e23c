"""GPU checks of the batched scan engine (ScanBatch).

* precision "fp64": every pair must equal the single-pair drop-in
  (itself checked against the oracle) to 1e-12 relative.
* precision "fp32": direct integrals stay FP64 (GEMV on the VALUs) and must
  match to 1e-10; exchange integrals come from the FP32 MFMA stacked GEMM and
  must satisfy |dI| <= 1e-6 * sum_ij |alpha_i| K_ij |beta_j| (the
  cancellation-aware bound: FP32 rounding of K and beta gives ~1e-8 of it).
* a captured HIP graph replays to the eager result, several times, with
  eager work and synchronisations in between.
"""
import numpy as np
import pytest

from oracle import radial as R, semisep
from dwba_amd import synthetic as S

pytestmark = pytest.mark.gpu


def _small_scan(seed=0, n_e=3, n_p=21, n_range=(1500, 3000), ionization=False):
    if ionization:
        return S.ionization_scan(n_points=n_e, n_triples=n_p, n_range=n_range, l_max=12, seed=seed)
    return S.excitation_scan(n_energies=n_e, n_pairs=n_p, n_range=n_range, l_max=15, k_range=(0.5, 4.0), seed=seed)


def _dropin(en, pr, L_max, **kw):
    from dwba_amd.dwba_matrix_elements import radial_ME_all_L_gpu
    bf = pr[2] if len(pr) > 2 else en.bound_f
    return radial_ME_all_L_gpu(en.grid, en.V_core, en.U_i, en.bound_i, bf, pr[0], pr[1], L_max, U_f_array=en.U_f, **kw)


@pytest.mark.parametrize("ionization", [False, True])
def test_fp64_batch_equals_dropin(ionization):
    from dwba_amd.batch import ScanBatch
    scan = _small_scan(1, ionization=ionization)
    b = ScanBatch(scan, 8, precision="fp64", oscillatory_method="full_split")
    b.run()
    res = b.integrals()
    mag = b.magnitudes()                   # sum |alpha| K |beta| per (pair, family, L)
    k = 0
    for ei, en in enumerate(scan):
        for pi, pr in enumerate(en.pairs):
            ref = _dropin(en, pr, 8, oscillatory_method="full_split")
            # same FP64 sums in a different order (the batch evaluates the shared
            # direct beta once per energy): rounding-level differences
            for L in range(9):
                for fam, a, r_ in ((0, res[ei][pi].I_L_direct[L], ref.I_L_direct[L]),
                                   (1, res[ei][pi].I_L_exchange[L], ref.I_L_exchange[L])):
                    assert abs(a - r_) <= 1e-12 * abs(r_) + 1e-14 * mag[k, fam, L], (ei, pi, L, a, r_)
            k += 1


@pytest.mark.parametrize("contraction", ["block", "dense"])
@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
@pytest.mark.parametrize("ionization", [False, True])
def test_mfma_batch_against_oracle(ionization, precision, contraction):
    """Exchange (and, for ionization, direct) through the stacked MFMA
    contraction, block form and dense tiles: |dI| <= 1e-6 sum|a K b| (FP32
    ~5e-8 of it, split-BF16x3 with dense tiles ~7e-7 on these short grids);
    split-BF16x3 in block form <= 3e-6 (measured ~1e-6: both operands of its
    three products carry ~16 significand bits, see test_bf16_split.py)."""
    from dwba_amd.batch import ScanBatch
    scan = _small_scan(2, n_e=2, n_p=40, ionization=ionization)
    L_max = 10
    b = ScanBatch(scan, L_max, precision=precision, oscillatory_method="advanced", contraction=contraction)
    b.run()
    res = b.results()
    tol = 3e-6 if (precision, contraction) == ("bf16x3", "block") else 1e-6
    k = 0
    worst = 0.0
    for en in scan:
        g = en.grid
        for pr in en.pairs:
            bf = pr[2] if len(pr) > 2 else en.bound_f
            red, parts = R.radial_ME_all_L_reduced(g.r, g.w_simpson, g.w_trapz, en.V_core, en.U_i, en.bound_i, bf,
                                                   pr[0], pr[1], L_max, U_f_array=en.U_f, return_parts=True)
            mag_e = semisep.abs_bound_all_L(g.r, parts["alpha_ex"], parts["beta_ex"], L_max)
            mag_d = semisep.abs_bound_all_L(g.r, parts["alpha_dir"], parts["beta_dir"], L_max)
            for L in range(L_max + 1):
                d_ref, e_ref = red.I_L_direct[L], red.I_L_exchange[L]
                if ionization:   # direct family also through the FP32 stacked GEMM
                    assert abs(res[k, 0, L] - d_ref) <= tol * mag_d[L] + 10 * parts["tail_noise"][L]
                else:
                    assert abs(res[k, 0, L] - d_ref) <= 1e-10 * abs(d_ref) + 1e-13 * mag_d[L] + 10 * parts["tail_noise"][L]
                err = abs(res[k, 1, L] - e_ref) / mag_e[L]
                worst = max(worst, err)
                assert err <= tol, (k, L, res[k, 1, L], e_ref, err)
            k += 1
    print(precision, contraction, "worst exchange |dI|/sum|aKb| =", worst)


def test_graph_replay_matches_eager():
    import torch
    from dwba_amd.batch import ScanBatch
    scan = _small_scan(3, n_e=2, n_p=30)
    b = ScanBatch(scan, 6, precision="fp32")
    b.launch()
    eager = b.results().copy()
    b.capture()
    for rep in range(3):
        b.out.zero_()
        torch.cuda.synchronize()
        b.run()
        torch.cuda.synchronize()
        assert np.array_equal(b.results(), eager), rep
        _ = (torch.ones(1000, device="cuda") * rep).sum().item()   # eager work between replays


def test_autotuned_batch_matches_default(tmp_path, monkeypatch):
    """Autotuning only reorders FP64 operations: the tuned batch equals the
    library-default one to rounding, and a second batch of the same
    signature takes the cached choice without timing."""
    from dwba_amd import autotune as AT
    from dwba_amd.batch import ScanBatch
    monkeypatch.setenv("DWBA_AUTOTUNE_CACHE", str(tmp_path / "tune.json"))
    tuner = AT.Autotuner()
    scan = _small_scan(4)
    ref = ScanBatch(scan, 10, precision="fp64", oscillatory_method="full_split")
    ref.run()
    r0 = ref.results()
    b = ScanBatch(scan, 10, precision="fp64", oscillatory_method="full_split")
    b.capture()
    ch = b.tune(tuner)
    print(f"\n[autotune] choice {ch.params} ({ch.ms:.3f} ms) of {len(ch.timings)} candidates", flush=True)
    assert not ch.cached
    assert len(ch.timings) == len(AT.semisep_candidates(b.L_max)) * (2 if b.ss_gemv is not None else 1)
    for _ in range(2):
        b.run()
    r1 = b.results()
    mag = ref.magnitudes()
    assert np.all(np.abs(r1 - r0) <= 1e-12 * np.abs(r0) + 1e-14 * mag)
    b2 = ScanBatch(scan, 10, precision="fp64", oscillatory_method="full_split")
    ch2 = b2.tune(AT.Autotuner())
    assert ch2.cached and ch2.params == ch.params


@pytest.mark.parametrize("L_max", [12, 20, 31])
def test_direct_gemv_path_matches_full_recurrence(L_max):
    """The shared-beta direct family (O(N) GEMV once per energy + a dot per
    pair, on the FP64 matrix cores or tiled on the VALUs) equals the per-pair
    recurrence to rounding.  70 pairs per energy leave a partial 16-pair item;
    L_max 12 uses one 16-multipole MFMA tile, 20 and 31 two."""
    from dwba_amd.batch import ScanBatch
    scan = _small_scan(6, n_e=2, n_p=70)
    b = ScanBatch(scan, L_max, precision="fp64", oscillatory_method="full_split")
    assert b.ss_gemv is not None
    b.run()
    r0 = b.results()
    mag = b.magnitudes()
    for kind in ("stage", "mfma", "tiled"):
        b.direct_gemv = True
        b.dot_kind = kind
        b.out.zero_()
        for st in b.E:             # prep rewrites only the beta_dir the shared path reads
            st.bd.zero_()
        b.run()
        r1 = b.results()
        excess = np.abs(r1 - r0) / (1e-12 * np.abs(r0) + 1e-14 * mag)
        w = np.unravel_index(np.argmax(excess), excess.shape)
        assert excess[w] <= 1.0, (kind, w, r1[w], r0[w], mag[w], int(np.sum(excess > 1)))
    # the lane-serial kernel on both job sets (exchange-only scans end at N1)
    b.semisep_lanes = True
    for gemv in (True, False):
        b.direct_gemv = gemv
        b.run()
        r2 = b.results()
        assert np.all(np.abs(r2 - r0) <= 1e-12 * np.abs(r0) + 1e-14 * mag), gemv


@pytest.mark.parametrize("mode", [None, "start", "after_prep"])
def test_tail_overlap_modes(mode):
    """The tail chain on a side stream (forked before or after prep, joined
    before finalize) gives the single-stream result exactly, eagerly and as
    replays of a captured graph with eager work in between."""
    import torch
    from dwba_amd.batch import ScanBatch
    scan = _small_scan(7, n_e=3, n_p=25)
    b = ScanBatch(scan, 10, precision="fp64", oscillatory_method="full_split")
    assert b.n_tail > 0
    b.overlap = None
    b.launch()
    ref = b.results().copy()
    b.overlap = mode
    b.out.zero_()
    b.tails.zero_()
    b.launch()
    assert np.array_equal(b.results(), ref)
    b.capture()
    for rep in range(3):
        b.out.zero_()
        b.tails.zero_()
        torch.cuda.synchronize()
        b.run()
        torch.cuda.synchronize()
        assert np.array_equal(b.results(), ref), (mode, rep)
        _ = (torch.ones(1000, device="cuda") * rep).sum().item()


def test_lane_tables_bit_identical():
    """The table-driven lane kernel reads 1/r, q and q^(lb b) computed once
    per grid with the same operations the lane kernel uses: identical bits,
    for both multipole block sizes and both job sets.  Its split-range form
    (two wavefronts per range, joined through LDS) agrees to rounding."""
    from dwba_amd.batch import ScanBatch
    scan = _small_scan(8, n_e=3, n_p=30)
    b = ScanBatch(scan, 30, precision="fp64", oscillatory_method="full_split")
    b.run()
    mag = b.magnitudes()
    b.semisep_lanes = True
    for lb in (4, 8):
        b.semisep_config = (0, lb)
        for gemv in (False, True):
            b.direct_gemv = gemv and b.ss_gemv is not None
            b.lane_tables = False
            b.out.zero_()
            b.run()
            r0 = b.results()
            b.lane_tables = True
            b.lane_split = 0
            b.out.zero_()
            b.run()
            r1 = b.results()
            assert np.array_equal(r0, r1), (lb, gemv)
            # jobs grouped by energy instead of sorted by length: same bits
            b.lane_order = "energy"
            b.out.zero_()
            b.run()
            assert np.array_equal(r0, b.results()), (lb, gemv, "energy order")
            b.lane_order = "length"
            # the range split between two wavefronts (4: LDS-staged, pairs of an energy padded to whole
            # groups): same sums, other order
            for occ in (2, 3, 4):
                b.lane_split = occ
                b.out.zero_()
                b.run()
                r2 = b.results()
                assert np.all(np.abs(r2 - r0) <= 1e-12 * np.abs(r0) + 1e-14 * mag), (lb, gemv, occ)
            b.lane_split = 0


@pytest.mark.parametrize("count,L_max,nmax", [(16, 12, 1000), (6, 12, 37), (16, 31, 515), (11, 20, 16), (16, 30, 3)])
def test_ddot_mfma_against_numpy(count, L_max, nmax):
    """dwba_ddot_mfma_f64 through the C ABI on one item of synthetic rows:
    out[p][L] = sum_{i < n_p} alpha_p[i] t[L][i], alpha zero from n_p to nmax."""
    import ctypes
    import torch
    from dwba_amd import _hip
    from dwba_amd.batch import DdotPair
    rng = np.random.default_rng(count * 1000 + L_max + nmax)
    nl = L_max + 1
    ld = ((nmax + 31) // 32) * 32
    ns = rng.integers(0, nmax + 1, size=count)
    ns[0] = nmax
    A = np.zeros((count, ld))
    for p in range(count):
        A[p, :ns[p]] = rng.standard_normal(ns[p])
    T = np.zeros((nl, ld))
    T[:, :nmax] = rng.standard_normal((nl, nmax))
    dev = torch.device("cuda", _hip.ensure_init(0))
    A_d = torch.as_tensor(A, device=dev)
    T_d = torch.as_tensor(T, device=dev)
    out = torch.full((count, 2, nl), np.nan, dtype=torch.float64, device=dev)
    pairs = (DdotPair * count)(*[DdotPair(_hip.ptr(A_d[p]), _hip.ptr(T_d), int(ns[p]), ld, p, 0)
                                 for p in range(count)])
    pairs_d = _hip.struct_array_to_device(pairs, dev)
    items = torch.as_tensor(np.asarray([[0, count]], dtype=np.int32), device=dev)
    lib = _hip.load_library()
    _hip.check(lib.dwba_ddot_mfma_f64(1, _hip.ptr(items), _hip.ptr(pairs_d), L_max, _hip.ptr(out),
                                      _hip.stream_handle()), "dwba_ddot_mfma_f64")
    got = out.cpu().numpy()[:, 0, :]
    ref = A[:, :nmax] @ T[:, :nmax].T
    scale = np.abs(A[:, :nmax]) @ np.abs(T[:, :nmax]).T
    bad = np.argwhere(~(np.abs(got - ref) <= 1e-14 * scale + 1e-300))
    assert bad.size == 0, ("wrong (pair, L) cells", bad.tolist()[:40], len(bad))


@pytest.mark.parametrize("counts,L_max,nmax", [((16,), 30, 1000), ((16,), 31, 37), ((11,), 12, 515), ((1,), 20, 16),
                                                ((16,), 30, 3), ((16,), 15, 32), ((5,), 8, 0),
                                                ((16, 13, 7, 16, 1, 10, 16, 12, 3), 30, 700)])
def test_ddot_stage_against_numpy(counts, L_max, nmax):
    """dwba_ddot_stage_f64 through the C ABI: items of up to 16 pairs, each
    with its own t, ragged n per pair (alpha zero from n_p to the item's
    largest n), chunk-boundary lengths, an empty item, nine items over the
    eight XCD runs: out[p][L] = sum_{i < n_p} alpha_p[i] t[L][i]."""
    import torch
    from dwba_amd import _hip
    from dwba_amd.batch import DdotPair
    rng = np.random.default_rng(sum(counts) * 1000 + L_max + nmax)
    nl = L_max + 1
    ld = max(32, ((nmax + 31) // 32) * 32)
    dev = torch.device("cuda", _hip.ensure_init(0))
    P = sum(counts)
    A = np.zeros((P, ld))
    Ts, pairs, items, refs, scales = [], [], [], [], []
    out = torch.full((P, 2, nl), np.nan, dtype=torch.float64, device=dev)
    A_rows = []
    first = 0
    for it, cnt in enumerate(counts):
        ns = rng.integers(0, nmax + 1, size=cnt)
        ns[0] = nmax
        T = np.zeros((nl, ld))
        T[:, :nmax] = rng.standard_normal((nl, nmax))
        for q in range(cnt):
            A[first + q, :ns[q]] = rng.standard_normal(ns[q])
        Ts.append((T, ns))
        items.append((first, cnt))
        first += cnt
    A_d = torch.as_tensor(A, device=dev)
    T_d = [torch.as_tensor(T, device=dev) for T, _ in Ts]
    first = 0
    for it, cnt in enumerate(counts):
        for q in range(cnt):
            pairs.append(DdotPair(_hip.ptr(A_d[first + q]), _hip.ptr(T_d[it]), int(Ts[it][1][q]), ld, first + q, 0))
        first += cnt
    pairs_d = _hip.struct_array_to_device((DdotPair * P)(*pairs), dev)
    items_d = torch.as_tensor(np.asarray(items, dtype=np.int32), device=dev)
    lib = _hip.load_library()
    print(f"\n[ddot stage] items {counts} L_max={L_max} nmax={nmax}", flush=True)
    _hip.check(lib.dwba_ddot_stage_f64(len(counts), _hip.ptr(items_d), _hip.ptr(pairs_d), L_max, _hip.ptr(out),
                                       _hip.stream_handle()), "dwba_ddot_stage_f64")
    got = out.cpu().numpy()[:, 0, :]
    first = 0
    for it, cnt in enumerate(counts):
        T = Ts[it][0]
        a = A[first:first + cnt, :nmax]
        ref = a @ T[:, :nmax].T
        scale = np.abs(a) @ np.abs(T[:, :nmax]).T
        g = got[first:first + cnt]
        bad = np.argwhere(~(np.abs(g - ref) <= 1e-14 * scale + 1e-300))
        assert bad.size == 0, ("wrong (pair, L) cells", it, bad.tolist()[:40], len(bad))
        first += cnt
    assert np.all(np.isnan(out.cpu().numpy()[:, 1, :]))       # the exchange slots are not touched


@pytest.mark.parametrize("split,lb", [(2, 8), (4, 8), (4, 4)])
@pytest.mark.parametrize("gemv", [False, True])
def test_pipelined_step_bit_identical(gemv, split, lb):
    """The two-half pipelined step (prep of the second half of the energies
    beside the contraction of the first, on two streams, tails on a third)
    gives the single-stream result bit for bit, eagerly and as graph replays;
    split 4: the LDS-staged lane kernel, whose groups never span energies."""
    import torch
    from dwba_amd.batch import ScanBatch
    scan = _small_scan(9, n_e=4, n_p=24)
    b = ScanBatch(scan, 20, precision="fp64", oscillatory_method="full_split")
    assert b._pipe is not None and 0 < b._pipe["kA"] < b.n_pairs
    b.semisep_config = (0, lb)
    b.semisep_lanes = True
    b.lane_split = split
    b.direct_gemv = gemv and b.ss_gemv is not None
    b.overlap = None
    b.pipeline = False
    b.launch()
    ref = b.results().copy()
    for mode in (None, "start", "after_prep"):
        b.overlap = mode
        b.pipeline = True
        assert b.pipeline_ok()
        b.out.zero_()
        b.launch()
        assert np.array_equal(b.results(), ref), mode
    b.capture()
    for rep in range(3):
        b.out.zero_()
        torch.cuda.synchronize()
        b.run()
        torch.cuda.synchronize()
        assert np.array_equal(b.results(), ref), rep
        _ = (torch.ones(1000, device="cuda") * rep).sum().item()


@pytest.mark.parametrize("ionization", [False, True])
def test_grouped_filon_prep_matches_per_pair(ionization):
    """Filon cell sums formed once per group of pairs (same energy, split
    index and layout) and applied to each pair by the fused pass give the
    per-pair prep's alpha / beta to rounding, hence the same integrals; the
    stacked vectors, the shared direct beta and the L = 0 scalars are cleared
    first so every value must be rewritten."""
    from dwba_amd.batch import ScanBatch
    scan = _small_scan(10, n_e=3, n_p=30, ionization=ionization)
    b = ScanBatch(scan, 12, precision="fp64", oscillatory_method="full_split")
    assert b.filon_groups is not None
    ng, n_members = b.filon_groups[1], len(b._grouped_pairs)
    print(f"\n[grouped prep] {ng} groups for {n_members} Filon pairs", flush=True)
    assert ng < n_members
    b.run()
    r0 = b.results()
    mag = b.magnitudes()
    vecs0 = [[m.clone() for m in (st.ad, st.ae, st.be)] for st in b.E]
    sc0 = b.scalars.clone()
    b.prep_grouped = True
    for st in b.E:
        for m in (st.ad, st.ae, st.be, st.bd):
            m.zero_()
    b.scalars.zero_()
    b.out.zero_()
    b.run()
    r1 = b.results()
    a, c = b.scalars.cpu().numpy(), sc0.cpu().numpy()
    assert np.all(np.abs(a - c) <= 1e-12 * np.abs(c) + 1e-15 * np.max(np.abs(c), axis=0) + 1e-300)
    for st, v0 in zip(b.E, vecs0):
        for m, m0 in zip((st.ad, st.ae, st.be), v0):
            a, c = m.cpu().numpy(), m0.cpu().numpy()
            tol = 1e-12 * np.abs(c) + 1e-15 * np.max(np.abs(c)) + 1e-300
            assert np.all(np.abs(a - c) <= tol), np.max(np.abs(a - c) / tol)
    assert np.all(np.abs(r1 - r0) <= 1e-12 * np.abs(r0) + 1e-14 * mag)


@pytest.mark.parametrize("ionization", [False, True])
def test_ionic_target_scan_matches_dropin(ionization):
    """Waves of an ionic target (Coulomb eta != 0, sigma_l): Filon phase
    nodes stepped on the device (generate_phase_nodes' Coulomb branch) and
    Coulomb phases in the outer tails.  The batch equals the single-pair
    drop-in to rounding in the default step, with grouped Filon prep (the
    group's cell pass reads the node count the stepping wrote) and in the
    two-half pipelined step, eagerly and as graph replays."""
    import torch
    from dwba_amd.batch import ScanBatch
    if ionization:
        scan = S.ionization_scan(n_points=2, n_triples=16, n_range=(1500, 2500), l_max=10, seed=13, z_ion=1.0)
    else:
        scan = S.excitation_scan(n_energies=3, n_pairs=16, n_range=(1500, 2500), l_max=12, k_range=(0.5, 3.0),
                                 seed=13, z_ion=1.0)
    L_max = 8
    b = ScanBatch(scan, L_max, precision="fp64", oscillatory_method="full_split")
    assert b.any_coul and b.filon_groups is not None
    refs = []
    for en in scan:
        for pr in en.pairs:
            d = _dropin(en, pr, L_max, oscillatory_method="full_split")
            refs.append([[d.I_L_direct[L] for L in range(L_max + 1)], [d.I_L_exchange[L] for L in range(L_max + 1)]])
    refs = np.asarray(refs)
    b.run()
    mag = b.magnitudes()
    tol = 1e-12 * np.abs(refs) + 1e-14 * mag

    def check(label):
        got = b.results()
        bad = np.argwhere(~(np.abs(got - refs) <= tol))
        assert bad.size == 0, (label, bad[:10].tolist(), got[tuple(bad[0])], refs[tuple(bad[0])])

    b.out.zero_()
    b.run()
    check("default")
    b.prep_grouped = True
    for st in b.E:
        for m in (st.ad, st.ae, st.be):
            m.zero_()
    b.out.zero_()
    b.run()
    check("grouped")
    b.prep_grouped = False
    b.semisep_lanes = True
    b.lane_split = 2
    b.pipeline = True
    assert b.pipeline_ok()
    b.out.zero_()
    b.run()
    check("pipelined")
    b.lane_split = 4                     # the LDS-staged lane kernel, 4 multipoles per lane
    b.semisep_config = (0, 4)
    b.out.zero_()
    b.run()
    check("pipelined, LDS-staged")
    b.capture()
    for rep in range(2):
        b.out.zero_()
        torch.cuda.synchronize()
        b.run()
        torch.cuda.synchronize()
        check(f"pipelined replay {rep}")
        _ = (torch.ones(1000, device="cuda") * rep).sum().item()
