"""Host planner (product) against the oracle's restatement: split index,
phase increments, phase-interval layout, reference tables (no GPU)."""
import numpy as np
import pytest

from oracle import grid as OG, quadrature as Q, radial as R, tail as T
from dwba_amd import planner, synthetic as S
from dwba_amd.planner import GridStats


@pytest.mark.parametrize("n", [2, 3, 5, 7, 9, 15])
def test_cc_reference_tables_bit_identical(n):
    x, w = planner.cc_reference(n)
    xo, wo = Q.cc_ref(n)
    assert np.array_equal(x, xo) and np.array_equal(w, wo)


def test_levin_tables_bit_identical():
    D, x = planner.levin_tables(8)
    assert np.array_equal(D, T.chebyshev_diff_matrix(8))
    assert np.array_equal(x, np.cos(np.pi * np.arange(8) / 7)[::-1])


def _waves(r, rng, with_match=True):
    ki, kf = rng.uniform(0.3, 4.0, 2)
    li, lf = rng.integers(0, 25, 2)
    return (S.make_wave(r, ki, li, rng.uniform(0, np.pi), with_match),
            S.make_wave(r, kf, lf, rng.uniform(0, np.pi), with_match))


@pytest.mark.parametrize("seed", range(12))
def test_split_index_matches_oracle(seed):
    rng = np.random.default_rng(seed)
    g = OG.make_r_grid(1e-4, 200.0, 3000)
    r, wt = g[0], g[1]
    ci, cf = _waves(r, rng, with_match=seed % 3 != 0)
    ui, uf = S.u1s(r), S.u2p(r) if seed % 2 else S.slater(r, 0.3)
    Ui = -np.exp(-r) / r * (1 + r)
    Uf = Ui * 0.7 if seed % 4 else None
    gs = GridStats.of(r, wt)
    got = planner.split_index(gs, Ui, Uf, ui, uf, ci.k_au, cf.k_au, ci.l, cf.l, ci.idx_match, cf.idx_match, 0.1)
    ref = R.compute_idx_limit(r, wt, Ui, Uf, ui, uf, ci.k_au, cf.k_au, ci.l, cf.l, ci.idx_match, cf.idx_match, 0.1)
    assert got == ref


@pytest.mark.parametrize("k,eta,n1", [(0.7, 0.0, 2500), (9.0, 0.0, 3000), (3.0, -1.5, 2000), (12.0, 0.8, 2999)])
def test_phase_increment_and_layout_match_oracle(k, eta, n1):
    r, wt, _ = OG.make_r_grid(1e-4, 200.0, 3000)
    gs = GridStats.of(r, wt)
    for req in (1.5708, 0.3, 0.05):
        got = planner.phase_increment(gs, n1, k, req, eta)
        ref = Q.adaptive_phase_increment(r[:n1], k, req, eta)
        assert got == ref
    lay = planner.filon_layout(gs, n1, k, 1.5708, eta)
    dphi = Q.adaptive_phase_increment(r[:n1], k, 1.5708, eta)
    ref_nodes = Q.generate_phase_nodes(r[0], r[n1 - 1], k, dphi, eta)
    if eta == 0.0:
        assert lay.mode == planner.PREP_FILON and not lay.coulomb
        assert lay.n_pnodes == len(ref_nodes)
        m = np.arange(lay.n_pnodes, dtype=np.float64)
        nodes = m * lay.p_step + lay.p_start
        nodes[-1] = lay.p_end
        assert np.array_equal(nodes, ref_nodes)
    else:
        assert lay.mode == planner.PREP_FILON and lay.coulomb and lay.p_step == dphi


def test_log_spacing_flags():
    r, wt, _ = OG.make_r_grid(1e-4, 200.0, 3000)
    gs = GridStats.of(r, wt)
    for n in (2, 3, 10, 3000):
        assert gs.log_spaced(n) == Q.is_log_spaced_grid(r[:n])
    lin = np.linspace(0.01, 10, 500)
    gl = GridStats.of(lin, OG.trapz_weights(lin))
    assert gl.log_spaced(500) == Q.is_log_spaced_grid(lin) == False


def test_filon_layout_degenerate_cases():
    r, wt, _ = OG.make_r_grid(1e-4, 200.0, 3000)
    gs = GridStats.of(r, wt)
    assert planner.filon_layout(gs, 1, 3.0, 1.5708, 0.0).mode == planner.PREP_ZERO
    assert planner.filon_layout(gs, 100, 1e-8, 1.5708, 0.0).mode == planner.PREP_STANDARD


def test_filon_limits_env(monkeypatch):
    monkeypatch.setenv("DWBA_MAX_FILON_SEGMENTS", "50")
    monkeypatch.setenv("DWBA_MAX_EFFECTIVE_DPHI", "0.01")
    seg, d = planner.filon_limits()
    assert seg == 128 and d == np.pi / 16
    assert T.filon_max_segments() == seg and T.filon_max_effective_dphi() == d


def test_filon_groups_share_geometry():
    """Grouping of Filon pairs for the shared cell sums: every member of a
    group has the same energy, split index, layout and interpolation kinds;
    no pair is in two groups; non-Filon pairs are in none; on a benchmark
    energy the ~400 pairs fall into a few dozen groups."""
    import numpy as np
    from dwba_amd import synthetic as S, planner as P
    from dwba_amd.dwba_matrix_elements import plan_pair, _fit
    scan = S.excitation_scan(n_energies=2, n_pairs=120, n_range=(4000, 6000), l_max=40, k_range=(0.5, 8.0), seed=3)
    plans = []
    for ei, en in enumerate(scan):
        r = np.asarray(en.grid.r)
        gs = P.GridStats.of(r, en.grid.w_trapz)
        N = len(r)
        Ui, ui, uf = _fit("U_i", en.U_i, N), _fit("u", en.bound_i.u_of_r, N), _fit("u", en.bound_f.u_of_r, N)
        ext = {}
        for pr in en.pairs:
            plans.append((ei, plan_pair(gs, Ui, None, ui, uf, pr[0], pr[1], 30, True, "full_split", 1.5708, 0.1, 0.5,
                                        True, ext)))
    groups = P.filon_groups(plans)
    seen = set()
    for g in groups:
        e0, p0 = plans[g[0]]
        for pos in g:
            e, p = plans[pos]
            assert pos not in seen
            seen.add(pos)
            assert e == e0 and p.N1 == p0.N1 and p.log1 == p0.log1 and p.log2 == p0.log2
            assert p.layout == p0.layout
    for pos, (e, p) in enumerate(plans):
        assert (pos in seen) == (p.mode == P.PREP_FILON and p.N1 >= 2)
    assert 0 < len(groups) < len(seen) // 2


def test_tail_sweep_nodes_follows_the_oracle_branch_choice():
    """planner.tail_sweep_nodes (tail job ordering and the tail roofline's work)
    takes the Filon / Levin / quad branch, the segment count and the node
    up-scaling the oracle's compute_outer_integral_oscillatory takes."""
    import math
    from oracle import tail as T
    rng = np.random.default_rng(3)
    for _ in range(200):
        ka, kb = rng.uniform(0.2, 8.0, 2)
        la, lb = (int(x) for x in rng.integers(0, 41, 2))
        ea, eb = (rng.uniform(-2, 0, 2) if rng.random() < 0.3 else (0.0, 0.0))
        rm = float(rng.uniform(5, 200))
        f, lv = planner.tail_sweep_nodes(ka, la, ea, kb, lb, eb, rm, 1000.0)
        pf = T.PhaseFunctions(ka, la, 0.0, ea, 0.0, kb, lb, 0.0, eb, 0.0)
        ef = el = 0
        mid = 0.5 * (rm + 1000.0)
        for sg in (-1, 1):
            om = abs(pf.dphi(sg, mid))
            if om < 1e-10:
                continue
            h = (np.pi / 4) / om
            if abs(pf.d2phi(sg, mid)) * h * h < 0.1:
                nseg = min(max(1, int(np.ceil((1000.0 - rm) / h))), T.filon_max_segments())
                eff = om * (1000.0 - rm) / nseg
                neff = 5
                if eff > T.filon_max_effective_dphi():
                    neff = min(15, max(5, 5 + 2 * (math.ceil(eff / T.filon_max_effective_dphi()) - 1)))
                ef += nseg * (neff - 1) + 1
            else:
                el += 8 * max(1, int(np.ceil((1000.0 - rm) / h)))
        assert (f, lv) == (ef, el), (ka, la, ea, kb, lb, eb, rm)
