"""GPU parity of the O(N) semiseparable contraction (dwba_il_semisep_f64)
against the oracle's long-double contraction and against the dense FP64
kernel on the same jobs.

Bar (FP64): |I_gpu - I_oracle| <= 1e-10 |I_oracle| + 1e-13 * sum|a_i K_ij b_j|,
the same as for the dense kernel."""
import numpy as np
import pytest

from oracle import semisep
from dwba_amd import grid as G
from dwba_amd import synthetic as S

pytestmark = pytest.mark.gpu


def _t(x):
    import torch
    return torch.as_tensor(np.ascontiguousarray(x, dtype=np.float64), device="cuda")


def _jobs(specs, seed=0):
    """specs: (n, n_rows, grid kind) -> host vectors and device job tuples."""
    rng = np.random.default_rng(seed)
    host, dev = [], []
    for (n, n_rows, kind) in specs:
        if n == 1:
            r, w = np.array([0.7]), np.ones(1)
        else:
            g = G.exp_grid(n, 1e-5, 1e3) if kind == "exp" else G.make_r_grid(1e-4, 200.0, n)
            r, w = g.r, (g.w_simpson if n > 2 else np.ones(n))
        vecs = [rng.standard_normal(n) * w for _ in range(4)]
        vecs[0][n_rows:] = 0.0
        vecs[2][n_rows:] = 0.0
        host.append((r, vecs, n_rows, n))
        dev.append((_t(r), *[_t(v) for v in vecs], n_rows, n))
    return host, dev


def _check(host, got, L_max, rtol=1e-10, atol_rel=1e-13):
    worst = 0.0
    for k, (r, vecs, n_rows, n) in enumerate(host):
        for fam, (a, b) in enumerate(((vecs[0], vecs[1]), (vecs[2], vecs[3]))):
            ref = semisep.il_all_L(r, a, b, L_max)
            mag = semisep.abs_bound_all_L(r, a, b, L_max)
            err = np.abs(got[k, fam] - ref)
            tol = rtol * np.abs(ref) + atol_rel * mag
            worst = max(worst, float(np.max(err / np.maximum(mag, 1e-300))))
            assert np.all(err <= tol), (k, fam, n_rows, n, np.max(err / np.maximum(tol, 1e-300)))
    return worst


SPECS = [(2, 2, "exp"), (3, 1, "exp"), (64, 64, "exp"), (65, 30, "exp"), (257, 257, "exp"), (1000, 0, "exp"),
         (1000, 999, "exp"), (3001, 3001, "ref"), (4097, 2000, "ref"), (1, 1, "exp")]


@pytest.mark.parametrize("L_max", [0, 7, 8, 30, 63])
def test_semisep_random_jobs(L_max):
    from dwba_amd import _hip
    host, dev = _jobs(SPECS, seed=L_max)
    print(f"\n[semisep] {len(SPECS)} jobs L_max={L_max}", flush=True)
    got = _hip.il_semisep_f64_batch(dev, L_max).cpu().numpy()
    w = _check(host, got, L_max)
    print(f"  worst |d|/sum|aKb| {w:.2e}", flush=True)


@pytest.mark.parametrize("lb", [0, 4])
@pytest.mark.parametrize("seg", [1, 3, 64, 100000])
def test_semisep_lane_split_independent(seg, lb):
    """The result does not depend on how the grid is cut into lanes and tasks
    or on the L-block width."""
    from dwba_amd import _hip
    host, dev = _jobs(SPECS, seed=5)
    print(f"\n[semisep] seg={seg} lb={lb:#x}", flush=True)
    got = _hip.il_semisep_f64_batch(dev, 20, seg=seg, lb=lb).cpu().numpy()
    _check(host, got, 20)


@pytest.mark.parametrize("lb", [0, 4])
@pytest.mark.parametrize("L_max", [0, 7, 8, 30, 63])
def test_semisep_lanes_random_jobs(L_max, lb):
    """Lane-serial kernel (one lane per job, family and multipole block)
    against the oracle on the same jobs: empty rows, one-point grid, rows
    shorter than the grid, L blocks that do not fill a power of two."""
    from dwba_amd import _hip
    host, dev = _jobs(SPECS, seed=100 + L_max)
    print(f"\n[semisep lanes] {len(SPECS)} jobs L_max={L_max} lb={lb}", flush=True)
    got = _hip.il_semisep_f64_batch(dev, L_max, lanes=True, lb=lb).cpu().numpy()
    w = _check(host, got, L_max)
    print(f"  worst |d|/sum|aKb| {w:.2e}", flush=True)


@pytest.mark.parametrize("split", [0, 2, 3])
@pytest.mark.parametrize("lb", [0, 4])
@pytest.mark.parametrize("L_max", [0, 7, 30, 63])
def test_semisep_lane_tables_random_jobs(L_max, lb, split):
    """Table-driven lane kernel, whole range per lane or split between two
    wavefronts (compiled for 2 or 3 waves per SIMD), jobs given in reverse
    order through the output index, against
    the oracle: empty rows, one- and two-point grids, rows shorter than the
    grid (beta-only ranges), split points at the ends."""
    from dwba_amd import _hip
    host, dev = _jobs(SPECS, seed=200 + L_max)
    print(f"\n[semisep lane tables] {len(SPECS)} jobs L_max={L_max} lb={lb} split={split}", flush=True)
    got = _hip.il_semisep_f64_batch(dev, L_max, lanes=True, lb=lb, tables=True, split=bool(split),
                                    waves_per_simd=split).cpu().numpy()
    w = _check(host, got, L_max)
    print(f"  worst |d|/sum|aKb| {w:.2e}", flush=True)


def _shared_jobs(grids, seed=0):
    """grids: (n, grid kind, [(n_cols, n_rows), ...]) -> jobs sharing each
    grid tensor (the LDS-staged kernel's one table per grid)."""
    rng = np.random.default_rng(seed)
    host, dev = [], []
    for (n, kind, ext) in grids:
        g = G.exp_grid(n, 1e-5, 1e3) if kind == "exp" else G.make_r_grid(1e-4, 200.0, n)
        r_d = _t(g.r)
        for (nc, nrow) in ext:
            r, w = g.r[:nc], g.w_simpson[:nc]
            vecs = [rng.standard_normal(nc) * w for _ in range(4)]
            vecs[0][nrow:] = 0.0
            vecs[2][nrow:] = 0.0
            host.append((r, vecs, nrow, nc))
            dev.append((r_d, *[_t(v) for v in vecs], nrow, nc))
    return host, dev


# ragged extents on shared grids: alpha rows shorter than, equal to and one
# short of the grid, chunk-boundary lengths (15, 16, 17, 33), one- and
# two-point grids, no alpha rows at all, a group of 16 that fills exactly
STAGE_GRIDS = [
    (1000, "exp", [(1000, 1000), (1000, 999), (1000, 0), (999, 500), (500, 500), (17, 17), (16, 3), (15, 15),
                   (33, 1), (2, 2), (1, 1), (1000, 640)]),
    (3001, "ref", [(3001 - 37 * k, 2900 - 61 * k) for k in range(16)]),
    (300, "exp", [(300, 300 - 13 * k) for k in range(21)]),
    (5, "exp", [(5, 5), (4, 2)]),
]


@pytest.mark.parametrize("lb,L_max", [(8, 16), (8, 20), (8, 30), (8, 31), (8, 32), (8, 63),
                                       (4, 4), (4, 7), (4, 8), (4, 15), (4, 30), (4, 31), (4, 32), (4, 63)])
def test_semisep_lane_stage_random_jobs(lb, L_max):
    """LDS-staged split lane kernel (dwba_il_semisep_lanes_stage_f64) against
    the oracle: jobs of several grids padded to whole groups per grid (in
    reverse order through the output index), ragged extents inside a group,
    multipole blocks past L_max, 32 to 4 pairs per group (lb 4, L_max 4 .. 63)."""
    from dwba_amd import _hip
    host, dev = _shared_jobs(STAGE_GRIDS, seed=300 + L_max)
    print(f"\n[semisep lane stage] {len(host)} jobs L_max={L_max} lb={lb}", flush=True)
    got = _hip.il_semisep_stage_f64_batch(dev, L_max, lb=lb).cpu().numpy()
    w = _check(host, got, L_max)
    print(f"  worst |d|/sum|aKb| {w:.2e}", flush=True)


def test_semisep_lane_stage_mixed_tables_read_nan():
    """A group whose streams are on different grid tables is not computed: its
    outputs read NaN (no silent mixing), the other groups are exact.  17 jobs
    unpadded, 16 streams per group: group 0 = direct family of jobs 0-15 (one
    grid), group 1 = direct of job 16 (another grid) + exchange of jobs 0-14,
    group 2 = exchange of jobs 15 and 16."""
    from dwba_amd import _hip
    host, dev = _shared_jobs([(400, "exp", [(400, 400 - 9 * k) for k in range(16)])], seed=3)
    h2, d2 = _jobs([(200, 150, "exp")], seed=4)
    print("\n[semisep lane stage] mixed tables in one group", flush=True)
    got = _hip.il_semisep_stage_f64_batch(dev + d2, 30, pad=False, lb=8).cpu().numpy()
    for k, (r, vecs, n_rows, n) in enumerate(host):
        ref = semisep.il_all_L(r, vecs[0], vecs[1], 30)
        mag = semisep.abs_bound_all_L(r, vecs[0], vecs[1], 30)
        assert np.all(np.abs(got[k, 0] - ref) <= 1e-10 * np.abs(ref) + 1e-13 * mag), k
    assert np.all(np.isnan(got[16, 0])) and np.all(np.isnan(got[:, 1]))


def test_semisep_lanes_matches_scan_kernel():
    from dwba_amd import _hip
    host, dev = _jobs(SPECS, seed=11)
    print("\n[semisep lanes] vs wavefront-scan kernel", flush=True)
    a = _hip.il_semisep_f64_batch(dev, 30, lanes=True).cpu().numpy()
    b = _hip.il_semisep_f64_batch(dev, 30).cpu().numpy()
    for k, (r, vecs, n_rows, n) in enumerate(host):
        for fam, (x, y) in enumerate(((vecs[0], vecs[1]), (vecs[2], vecs[3]))):
            mag = semisep.abs_bound_all_L(r, x, y, 30)
            assert np.all(np.abs(a[k, fam] - b[k, fam]) <= 1e-12 * np.abs(b[k, fam]) + 1e-13 * mag)


def test_semisep_kernel_only_workload_n15000():
    from dwba_amd import _hip
    g, (ad, bd, ae, be, _) = S.kernel_only_workload(15000, seed=0)
    print("\n[semisep] kernel-only N=15000 L=0..30", flush=True)
    got = _hip.il_semisep_f64_batch([(_t(g.r), _t(ad), _t(bd), _t(ae), _t(be), 15000, 15000)], 30).cpu().numpy()
    _check([(g.r, [ad, bd, ae, be], 15000, 15000)], got, 30)


def test_semisep_matches_dense_kernel():
    from dwba_amd import _hip
    host, dev = _jobs(SPECS[2:], seed=9)
    print("\n[semisep] vs dense plain batch", flush=True)
    a = _hip.il_semisep_f64_batch(dev, 30).cpu().numpy()
    b = _hip.il_plain_f64_batch(dev, 30).cpu().numpy()
    for k, (r, vecs, n_rows, n) in enumerate(host):
        for fam, (x, y) in enumerate(((vecs[0], vecs[1]), (vecs[2], vecs[3]))):
            mag = semisep.abs_bound_all_L(r, x, y, 30)
            assert np.all(np.abs(a[k, fam] - b[k, fam]) <= 1e-12 * np.abs(b[k, fam]) + 1e-13 * mag)


def test_semisep_deterministic():
    from dwba_amd import _hip
    _, dev = _jobs(SPECS, seed=2)
    a = _hip.il_semisep_f64_batch(dev, 30).cpu().numpy()
    b = _hip.il_semisep_f64_batch(dev, 30).cpu().numpy()
    assert np.array_equal(a, b)


def test_semisep_rejects_bad_shapes():
    from dwba_amd import _hip
    g = G.exp_grid(100)
    r = _t(g.r)
    with pytest.raises(ValueError):
        _hip.il_semisep_f64_batch([(r, r, r, r, r, 101, 100)], 4)
    with pytest.raises(ValueError):
        _hip.il_semisep_f64_batch([(r, r, r, r, r, 10, 100)], 64)


@pytest.mark.parametrize("L_max", [0, 9, 30])
def test_semisep_gemv_matches_dense_gemv(L_max):
    """t_L = K_L beta per energy: O(N) recurrences against the dense FP64 GEMV
    kernel and the oracle's long-double K_L x."""
    import torch
    from dwba_amd import _hip
    from dwba_amd.batch import DgemvEnergy
    from oracle import gradient as OGr
    rng = np.random.default_rng(L_max)
    lib = _hip.load_library()
    _hip.ensure_init()
    specs = [(3000, 1800), (1201, 1201), (4500, 10), (2, 2)]
    grids, descs_a, descs_b, outs = [], [], [], []
    for n, rows in specs:
        g = G.exp_grid(n, 1e-5, 1e3)
        beta = rng.standard_normal(n) * g.w_simpson
        ld = (n + 31) // 32 * 32
        ta = torch.zeros((L_max + 1, ld), dtype=torch.float64, device="cuda")
        tb = torch.zeros_like(ta)
        r_d, b_d = _t(g.r), _t(beta)
        grids.append((g.r, beta, rows, r_d, b_d))
        descs_a.append(DgemvEnergy(_hip.ptr(r_d), _hip.ptr(b_d), _hip.ptr(ta), n, rows, ld, L_max))
        descs_b.append(DgemvEnergy(_hip.ptr(r_d), _hip.ptr(b_d), _hip.ptr(tb), n, rows, ld, L_max))
        outs.append((ta, tb))
    da = _hip.struct_array_to_device((DgemvEnergy * len(descs_a))(*descs_a), torch.device("cuda"))
    db = _hip.struct_array_to_device((DgemvEnergy * len(descs_b))(*descs_b), torch.device("cuda"))
    print(f"\n[semisep gemv] L_max={L_max}", flush=True)
    _hip.check(lib.dwba_semisep_gemv_f64(len(specs), _hip.ptr(da), L_max, _hip.stream_handle()), "gemv")
    _hip.check(lib.dwba_dgemv_f64(len(specs), _hip.ptr(db), max(r for _, r in specs), L_max, _hip.stream_handle()),
               "dgemv")
    torch.cuda.synchronize()
    for (r, beta, rows, *_), (ta, tb) in zip(grids, outs):
        a = ta.cpu().numpy()[:, :rows]
        b = tb.cpu().numpy()[:, :rows]
        ref = OGr.kmatvec_all_L(r, beta, L_max).astype(np.float64)[:, :rows]
        mag = OGr.kmatvec_all_L(r, np.abs(beta), L_max).astype(np.float64)[:, :rows]
        assert np.all(np.abs(a - ref) <= 1e-13 * mag + 1e-300), np.max(np.abs(a - ref) / mag)
        assert np.all(np.abs(b - ref) <= 1e-13 * mag + 1e-300)
