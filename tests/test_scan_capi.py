"""The whole-scan C entry points (dwba_scan_plan / _launch / _capture /
_replay / _results / _destroy, and the host-only dwba_scan_plan_pairs),
driven through ctypes alone: the inputs are plain numpy arrays in the
C structs of include/dwba_hip.h, as a non-Python host would pass them.

* CPU: the C++ planner makes the same per-pair decisions as the Python one
  (plan_pair: split index, prep mode, tail kind, extents, log-spacing tests,
  Filon phase layout) on excitation, ionization and ionic-target scans, on
  exponential and on the reference's linear-log grids, for every oscillatory
  method and with the oscillatory branch off.
* GPU: the integrals of a scan equal, pair by pair, the CPU oracle's
  restatement of radial_ME_all_L (to 1e-10 |I| + 1e-13 sum|a K b| + the
  oracle's own tail-quadrature noise) and the single-pair drop-in (FP64 sums
  in another order: 1e-12 |I| + 1e-14 sum|a K b|); a captured graph replays
  to the eager launch's bits.
"""
import ctypes
from dataclasses import dataclass

import numpy as np
import pytest

from dwba_amd import _hip
from dwba_amd import synthetic as S
from dwba_amd.grid import exp_grid, make_r_grid


@dataclass
class Point:
    """One energy of a scan as the reference's loops hold it."""
    grid: object
    V_core: np.ndarray
    U_i: np.ndarray
    U_f: np.ndarray
    bound_i: object
    pairs: list            # (cont_i, cont_f, bound_f or ejected wave)


def make_points(kind="excitation", n_energies=2, n_pairs=12, n_range=(1500, 2500), l_max=10, seed=0, z_ion=0.0,
                k_range=(0.5, 3.0), grid="exp"):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_energies):
        n = int(rng.integers(n_range[0], n_range[1] + 1))
        g = exp_grid(n, 1e-5, 1e3) if grid == "exp" else make_r_grid(1e-5, 200.0, n)
        r = g.r
        V, U = S.h_potentials(r)
        pairs = []
        if kind == "excitation":
            ki, kf = rng.uniform(*k_range, size=2)
            bi, bf = S.Bound(S.u1s(r)), S.Bound(S.u2p(r), l=1)
            while len(pairs) < n_pairs:
                li, lf = rng.integers(0, l_max + 1, size=2)
                di, df = rng.uniform(0.0, np.pi, size=2)
                pairs.append((S.make_wave(r, ki, li, di, z_ion=z_ion), S.make_wave(r, kf, lf, df, z_ion=z_ion), bf))
        else:
            bi = S.Bound(S.slater(r, 0.6))
            ke = rng.uniform(0.2, 2.0)
            k_i = 2.0
            kf = np.sqrt(max(k_i ** 2 - ke ** 2 - 0.36, 0.04))
            while len(pairs) < n_pairs:
                li, lf, le = rng.integers(0, l_max + 1, size=3)
                if (li + lf + le) % 2:
                    continue
                di, df, de = rng.uniform(0.0, np.pi, size=3)
                pairs.append((S.make_wave(r, k_i, li, di, z_ion=z_ion), S.make_wave(r, kf, lf, df, z_ion=z_ion),
                              S.make_wave(r, ke, le, de, z_ion=z_ion)))
        out.append(Point(g, V, U, U, bi, pairs))
    return out


def c_inputs(points):
    """dwba_scan_energy / _wave / _pair arrays (host pointers into numpy
    arrays kept alive by the returned list)."""
    keep = []

    def arr(a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        keep.append(a)
        return a.ctypes.data, len(a)

    E = (_hip.ScanEnergy * len(points))()
    waves, wave_id, pairs = [], {}, []
    uf_ptr = {}
    for e, p in enumerate(points):
        g = p.grid
        n = len(g.r)
        E[e] = _hip.ScanEnergy(arr(g.r)[0], arr(g.w_simpson)[0], arr(g.w_trapz)[0], arr(p.V_core)[0],
                               arr(p.U_i)[0], arr(p.U_f)[0] if p.U_f is not None else None, arr(p.bound_i.u_of_r)[0],
                               n, 0)
        for ci, cf, bf in p.pairs:
            ids = []
            for w in (ci, cf):
                if id(w) not in wave_id:
                    ptr, m = arr(w.chi_of_r)
                    wave_id[id(w)] = len(waves)
                    waves.append(_hip.ScanWave(ptr, w.k_au, w.phase_shift, w.eta, w.sigma_l, w.l, w.idx_match, m, 0))
                ids.append(wave_id[id(w)])
            if id(bf) not in uf_ptr:
                uf_ptr[id(bf)] = arr(bf.u_of_r)
            ptr, m = uf_ptr[id(bf)]
            pairs.append(_hip.ScanPair(ptr, m, e, ids[0], ids[1], int(not hasattr(bf, "k_au")), 0))
    W = (_hip.ScanWave * len(waves))(*waves)
    P = (_hip.ScanPair * len(pairs))(*pairs)
    return E, W, P, keep


def options(L_max, method="full_split", use_osc=True, CC_nodes=5, phase_increment=1.5708, mgf=0.10, k_thr=0.5):
    return _hip.ScanOptions(L_max, int(use_osc), _hip.SCAN_METHODS[method], CC_nodes, phase_increment, mgf, k_thr)


def python_plans(points, L_max, method="full_split", use_osc=True, phase_increment=1.5708, mgf=0.10, k_thr=0.5):
    from dwba_amd.dwba_matrix_elements import plan_pair
    from dwba_amd.planner import GridStats
    plans = []
    for p in points:
        gs = GridStats.of(p.grid.r, p.grid.w_trapz)
        ext = {}
        for ci, cf, bf in p.pairs:
            plans.append(plan_pair(gs, p.U_i, p.U_f, p.bound_i.u_of_r, bf.u_of_r, ci, cf, L_max, use_osc, method,
                                   phase_increment, mgf, k_thr, not hasattr(bf, "k_au"), ext))
    return plans


PLAN_CASES = [
    dict(kind="excitation", seed=1),
    dict(kind="excitation", seed=2, z_ion=1.0),
    dict(kind="ionization", seed=3),
    dict(kind="ionization", seed=4, z_ion=1.0),
    dict(kind="excitation", seed=5, grid="linlog"),
    dict(kind="excitation", seed=6, k_range=(0.05, 0.4)),            # below k_threshold: plain branch
    dict(kind="excitation", seed=7, method="legacy"),
    dict(kind="excitation", seed=8, method="advanced", use_osc=False),
    dict(kind="excitation", seed=9, n_energies=4, n_pairs=40, n_range=(3000, 6000), l_max=40, k_range=(0.5, 8.0)),
]


@pytest.mark.parametrize("case", PLAN_CASES, ids=lambda c: "-".join(f"{k}={v}" for k, v in c.items()))
def test_c_planner_matches_python(case):
    case = dict(case)
    method = case.pop("method", "full_split")
    use_osc = case.pop("use_osc", True)
    L_max = 12
    points = make_points(**case)
    lib = _hip.load_library()
    E, W, P, keep = c_inputs(points)
    out = (_hip.ScanPairPlan * len(P))()
    rc = lib.dwba_scan_plan_pairs(len(E), E, len(W), W, len(P), P, ctypes.byref(options(L_max, method, use_osc)), out)
    assert rc == 0, _hip.last_error()
    ref = python_plans(points, L_max, method, use_osc)
    modes = set()
    for k, (c, q) in enumerate(zip(out, ref)):
        lay = q.layout
        got = (c.N1, c.mode, c.tail, c.rows, c.cols, c.log1, c.log2, c.n_pnodes, c.coulomb, c.oscillatory)
        want = (q.N1, q.mode, q.tail_kind, q.rows, q.cols, q.log1, q.log2, lay.n_pnodes, int(lay.coulomb),
                int(q.oscillatory))
        assert got == want, (k, got, want)
        assert c.r_match == q.r_match
        # phase layout: the same formulas; the Coulomb phase-step maximum may
        # differ in the last bit (sum order), so compare to rounding
        for a, b in ((c.p_start, lay.p_start), (c.p_step, lay.p_step), (c.p_end, lay.p_end)):
            assert abs(a - b) <= 4e-16 * abs(b), (k, a, b)
        modes.add(c.mode)
    if case.get("k_range", (0.5, 3.0))[1] < 0.5 or not use_osc:
        assert modes == {0}
    else:
        assert 1 in modes


def test_c_planner_rejects_bad_input():
    lib = _hip.load_library()
    points = make_points(n_energies=1, n_pairs=2)
    E, W, P, keep = c_inputs(points)
    out = (_hip.ScanPairPlan * len(P))()
    P[1].wave_f = len(W)                               # out of range
    assert lib.dwba_scan_plan_pairs(len(E), E, len(W), W, len(P), P, ctypes.byref(options(8)), out) == 1
    assert b"wave index out of range" in _hip.last_error().encode()
    E, W, P, keep = c_inputs(points)
    W[0].n = len(points[0].grid.r) - 1                 # shorter than the grid (radial_ME_all_L_gpu rejects it)
    assert lib.dwba_scan_plan_pairs(len(E), E, len(W), W, len(P), P, ctypes.byref(options(8)), out) == 1
    assert b"shorter than the grid" in _hip.last_error().encode()
    E, W, P, keep = c_inputs(points)
    assert lib.dwba_scan_plan_pairs(len(E), E, len(W), W, len(P), P, ctypes.byref(options(64)), out) == 1


# ---------------------------------------------------------------- GPU
def _run_scan(points, L_max, method="full_split", capture=False):
    import torch
    lib = _hip.load_library()
    _hip.ensure_init()
    E, W, P, keep = c_inputs(points)
    h = ctypes.c_void_p()
    rc = lib.dwba_scan_plan(len(E), E, len(W), W, len(P), P, ctypes.byref(options(L_max, method)), ctypes.byref(h))
    assert rc == 0, _hip.last_error()
    info = (ctypes.c_int64 * 8)()
    assert lib.dwba_scan_info(h, info, 8) == 0
    nl = L_max + 1
    try:
        s = torch.cuda.Stream()
        sh = ctypes.c_void_p(s.cuda_stream)
        assert lib.dwba_scan_results(h)                  # device pointer, for device consumers
        host = np.empty((len(P), 2, nl))

        def fetch():
            assert lib.dwba_scan_read_results(h, host.ctypes.data, sh) == 0, _hip.last_error()
            return host.copy()

        assert lib.dwba_scan_launch(h, sh) == 0, _hip.last_error()
        eager = fetch()
        assert lib.dwba_scan_launch(h, sh) == 0, _hip.last_error()      # relaunch: same buffers, same bits
        again = fetch()
        assert np.array_equal(eager, again, equal_nan=True)
        if capture:
            assert lib.dwba_scan_capture(h, sh) == 0, _hip.last_error()
            for _ in range(3):
                assert lib.dwba_scan_replay(h, sh) == 0, _hip.last_error()
                rep = fetch()
                assert np.array_equal(eager, rep, equal_nan=True)
                torch.zeros(1 << 20, device="cuda").sum().item()      # eager work between replays
    finally:
        assert lib.dwba_scan_destroy(h) == 0
    return eager, list(info)


def _references(points, L_max, method="full_split", oracle_stride=1):
    """Drop-in per pair (GPU) and the oracle's reduced restatement (CPU) on
    every oracle_stride-th pair, with sum|a K b| per family and L."""
    from dwba_amd.dwba_matrix_elements import radial_ME_all_L_gpu
    from oracle import radial as R
    from oracle import semisep
    drop, orc = [], {}
    k = 0
    for p in points:
        g = p.grid
        for ci, cf, bf in p.pairs:
            d = radial_ME_all_L_gpu(g, p.V_core, p.U_i, p.bound_i, bf, ci, cf, L_max, oscillatory_method=method,
                                    U_f_array=p.U_f)
            drop.append([[d.I_L_direct[L] for L in range(L_max + 1)], [d.I_L_exchange[L] for L in range(L_max + 1)]])
            red, parts = R.radial_ME_all_L_reduced(g.r, g.w_simpson, g.w_trapz, p.V_core, p.U_i, p.bound_i, bf, ci, cf,
                                                   L_max, oscillatory_method=method, U_f_array=p.U_f,
                                                   return_parts=True)
            mag = np.array([semisep.abs_bound_all_L(g.r, parts["alpha_dir"], parts["beta_dir"], L_max),
                            semisep.abs_bound_all_L(g.r, parts["alpha_ex"], parts["beta_ex"], L_max)])
            I = np.array([[red.I_L_direct[L] for L in range(L_max + 1)],
                          [red.I_L_exchange[L] for L in range(L_max + 1)]])
            noise = np.array([parts["tail_noise"], np.zeros(L_max + 1)])
            orc[k] = (I, mag, noise)
            k += 1
    return np.asarray(drop), orc


def _check(got, drop, orc):
    for k, (I, mag, noise) in orc.items():
        bad = ~(np.abs(got[k] - I) <= 1e-10 * np.abs(I) + 1e-13 * mag + 10 * noise)
        assert not bad.any(), ("oracle", k, np.argwhere(bad)[:5].tolist(), got[k][bad][:3], I[bad][:3])
        bad = ~(np.abs(got[k] - drop[k]) <= 1e-12 * np.abs(drop[k]) + 1e-14 * mag)
        assert not bad.any(), ("drop-in", k, np.argwhere(bad)[:5].tolist(), got[k][bad][:3], drop[k][bad][:3])


GPU_CASES = [
    dict(kind="excitation", seed=21),
    dict(kind="excitation", seed=22, z_ion=1.0),
    dict(kind="ionization", seed=23),
    dict(kind="excitation", seed=24, grid="linlog", n_energies=1, method="legacy"),
]


@pytest.mark.gpu
@pytest.mark.parametrize("case", GPU_CASES, ids=lambda c: "-".join(f"{k}={v}" for k, v in c.items()))
def test_c_scan_matches_oracle_and_dropin(case):
    case = dict(case)
    method = case.pop("method", "full_split")
    L_max = 8
    points = make_points(**case)
    print("scan case", case, method, flush=True)
    got, info = _run_scan(points, L_max, method, capture=(case.get("seed") == 21))
    n_pairs = sum(len(p.pairs) for p in points)
    assert info[0] == n_pairs and info[1] == L_max
    if case["kind"] == "excitation" and method != "legacy":
        assert info[3] > 0                       # outer tails planned
        assert info[5] == len(points)            # one shared direct beta per energy
    if case.get("z_ion"):
        assert 0 < info[6] < n_pairs             # Coulomb node buffers shared within groups
    drop, orc = _references(points, L_max, method)
    _check(got, drop, orc)


@pytest.mark.gpu
@pytest.mark.parametrize("z_ion", [0.0, 1.0])
def test_cscan_binding_matches_scanbatch_step(z_ion):
    """The Python binding of the C engine (dwba_amd.cscan) on a
    multi-energy scan agrees with ScanBatch set to the same FP64 step (tail
    fork, shared direct beta with matrix-core dots, LDS-staged lane
    contraction at 4 multipoles per lane, grouped Filon prep) to FP64
    rounding."""
    from dwba_amd.batch import ScanBatch
    from dwba_amd.cscan import CScan
    scan = S.excitation_scan(n_energies=3, n_pairs=24, n_range=(1500, 3000), l_max=14, k_range=(0.5, 4.0), seed=31,
                             z_ion=z_ion)
    L_max = 12
    b = ScanBatch(scan, L_max, precision="fp64", oscillatory_method="full_split")
    b.overlap = "start"
    b.direct_gemv = True
    b.dot_kind = "stage"
    b.semisep_lanes = True
    b.lane_tables = True
    b.lane_split = 4
    b.semisep_config = (0, 4)
    b.lane_order = "energy"
    b.prep_grouped = True
    b.run()
    ref = b.results()
    mag = b.magnitudes()
    cs = CScan(scan, L_max, oscillatory_method="full_split")
    cs.launch()
    got = cs.results()
    info = cs.info()
    cs.close()
    assert info["pairs"] == ref.shape[0] and info["shared_beta_energies"] == len(scan)
    bad = ~(np.abs(got - ref) <= 1e-12 * np.abs(ref) + 1e-14 * mag)
    assert not bad.any(), (np.argwhere(bad)[:5].tolist(), got[bad][:3], ref[bad][:3])


@pytest.mark.gpu
def test_nonfinite_density_beyond_split_index():
    """A density that is not finite beyond idx_limit (where w_lim = 0) turns
    the L = 0 correction sums into NaN, as the reference's dot products do,
    so the correction is skipped for the pairs that use it; pairs whose
    densities are finite there may stop their prep at idx_limit.  The C scan
    and ScanBatch (grouped Filon prep, shared direct beta) agree with the
    drop-in on both kinds of pairs."""
    from dwba_amd.batch import ScanBatch, EnergyInput
    from dwba_amd.cscan import CScan
    from dwba_amd.dwba_matrix_elements import radial_ME_all_L_gpu
    points = make_points(kind="excitation", n_energies=1, n_pairs=10, seed=41)
    p = points[0]
    ci0 = p.pairs[0][0]
    bad = np.array(ci0.chi_of_r, dtype=np.float64)
    bad[len(bad) - 3] = np.inf                       # beyond every pair's split index
    ci0.chi_of_r = bad
    L_max = 6
    scan = [EnergyInput(grid=p.grid, V_core=p.V_core, U_i=p.U_i, bound_i=p.bound_i, bound_f=p.pairs[0][2],
                        pairs=[(ci, cf) for ci, cf, _ in p.pairs], U_f=p.U_f)]
    ref = []
    for ci, cf, bf in p.pairs:
        d = radial_ME_all_L_gpu(p.grid, p.V_core, p.U_i, p.bound_i, bf, ci, cf, L_max, oscillatory_method="full_split",
                                U_f_array=p.U_f)
        ref.append([[d.I_L_direct[L] for L in range(L_max + 1)], [d.I_L_exchange[L] for L in range(L_max + 1)]])
    ref = np.asarray(ref)
    b = ScanBatch(scan, L_max, precision="fp64", oscillatory_method="full_split")
    b.direct_gemv = b.ss_gemv is not None
    b.semisep_lanes = True
    b.lane_split = 2
    b.prep_grouped = True
    b.run()
    got_b = b.results()
    mag = b.magnitudes()
    cs = CScan(scan, L_max, oscillatory_method="full_split")
    cs.launch()
    got_c = cs.results()
    cs.close()
    uses_bad = np.array([ci is ci0 for ci, _, _ in p.pairs])
    assert uses_bad.any() and not uses_bad.all()
    # the 1s -> 2p overlap is far above 1e-12, so the NaN dot product reaches
    # I_0 (direct, and exchange where its overlap sum passes the same test) of
    # exactly the pairs that use the planted wave; every L >= 1 stays finite
    assert np.array_equal(np.isnan(ref[:, 0, 0]), uses_bad)
    assert not np.isnan(ref[~uses_bad]).any()
    assert np.all(np.isfinite(ref[:, :, 1:]))
    for got in (got_b, got_c):
        assert np.array_equal(np.isnan(got), np.isnan(ref))
        ok = np.isnan(ref) | (np.abs(got - ref) <= 1e-12 * np.abs(ref) + 1e-14 * mag)
        assert ok.all(), np.argwhere(~ok)[:5].tolist()
