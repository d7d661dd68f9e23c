"""GPU parity of il_plain_f64 against the oracle's exact O(N) contraction.

Bar: FP64 kernel, |I_gpu - I_oracle| <= 1e-10 |I_oracle| + 1e-13 * sum|a_i K_ij b_j|
(the second term only matters for integrals cancelled to below 1e-3 of their
magnitude).  The oracle evaluates in long double.
"""
import numpy as np
import pytest

from oracle import semisep
from dwba_amd import grid as G
from dwba_amd import synthetic as S

pytestmark = pytest.mark.gpu


def _t(x):
    import torch
    return torch.as_tensor(np.ascontiguousarray(x, dtype=np.float64), device="cuda")


def _run(r, ad, bd, ae, be, n_rows, n_cols, L_max):
    from dwba_amd import _hip
    out = _hip.il_plain_f64(_t(r), _t(ad), _t(bd), _t(ae), _t(be), n_rows, n_cols, L_max)
    return out.cpu().numpy()


def _check(r, ad, bd, ae, be, n_rows, n_cols, L_max, rtol=1e-10, atol_rel=1e-13):
    got = _run(r, ad, bd, ae, be, n_rows, n_cols, L_max)
    rr = r[:n_cols]
    a_d = np.zeros(n_cols); a_d[:n_rows] = ad[:n_rows]
    a_e = np.zeros(n_cols); a_e[:n_rows] = ae[:n_rows]
    for fam, (a, b) in enumerate(((a_d, bd[:n_cols]), (a_e, be[:n_cols]))):
        ref = semisep.il_all_L(rr, a, b, L_max)
        mag = semisep.abs_bound_all_L(rr, a, b, L_max)
        err = np.abs(got[fam] - ref)
        tol = rtol * np.abs(ref) + atol_rel * mag
        assert np.all(err <= tol), (fam, np.max(err / np.maximum(tol, 1e-300)), got[fam][:4], ref[:4])
    return got


@pytest.mark.parametrize("n,L_max", [(2, 0), (3, 4), (64, 4), (65, 8), (257, 30), (1000, 20), (3001, 30)])
def test_plain_random_square(n, L_max):
    rng = np.random.default_rng(n + L_max)
    g = G.exp_grid(n, 1e-4, 200.0)
    ad, bd, ae, be = (rng.standard_normal(n) * g.w_simpson for _ in range(4))
    _check(g.r, ad, bd, ae, be, n, n, L_max)


@pytest.mark.parametrize("n_rows,n_cols", [(0, 100), (1, 100), (63, 200), (700, 1500), (1499, 1500)])
def test_plain_rectangular(n_rows, n_cols):
    rng = np.random.default_rng(n_rows)
    g = G.exp_grid(n_cols, 1e-5, 1e3)
    ad, bd, ae, be = (rng.standard_normal(n_cols) for _ in range(4))
    ad[n_rows:] = 0.0
    ae[n_rows:] = 0.0
    _check(g.r, ad, bd, ae, be, n_rows, n_cols, 12)


def test_plain_kernel_only_workload_n15000():
    """Single pair, N = 15000 exp grid on [1e-5, 1000], L = 0..30 (hydrogenic
    1s 2p bound pair, seeded continuum waves)."""
    g, (ad, bd, ae, be, _) = S.kernel_only_workload(15000, seed=0)
    _check(g.r, ad, bd, ae, be, 15000, 15000, 30)


def test_plain_deterministic():
    g, (ad, bd, ae, be, _) = S.kernel_only_workload(4000, seed=3)
    a = _run(g.r, ad, bd, ae, be, 4000, 4000, 30)
    b = _run(g.r, ad, bd, ae, be, 4000, 4000, 30)
    assert np.array_equal(a, b)


def test_plain_batch_matches_single():
    import torch
    from dwba_amd import _hip
    jobs, singles = [], []
    for k, n in enumerate((500, 1203, 64, 3000)):
        g = G.exp_grid(n, 1e-5, 1e3)
        rng = np.random.default_rng(k)
        vecs = [rng.standard_normal(n) * g.w_simpson for _ in range(4)]
        n_rows = n if k % 2 == 0 else n // 2
        vecs[0][n_rows:] = 0.0
        vecs[2][n_rows:] = 0.0
        jobs.append((_t(g.r), *[_t(v) for v in vecs], n_rows, n))
        singles.append(_run(g.r, *vecs, n_rows, n, 20))
    out = _hip.il_plain_f64_batch(jobs, 20).cpu().numpy()
    for k in range(len(jobs)):
        np.testing.assert_allclose(out[k], singles[k], rtol=1e-13, atol=0)


def test_plain_rejects_bad_shapes():
    from dwba_amd import _hip
    g = G.exp_grid(100)
    with pytest.raises(ValueError):
        _run(g.r, g.r, g.r, g.r, g.r, 101, 100, 4)
    with pytest.raises(ValueError):
        _run(g.r, g.r, g.r, g.r, g.r, 10, 100, 64)
