"""dwba_il_plain_f64_noalloc: the workspace-free form of dwba_il_plain_f64.

Called through the C ABI twice in a row on the same GPU (the second call on
a smaller grid and fewer multipoles, so the shared scratch array still holds
the first call's larger plan), each result compared bit for bit with
dwba_il_plain_f64 on a caller-allocated workspace and with the oracle's
long-double O(N) contraction to 1e-10 relative.  Argument errors are
reported as status codes.
"""
import ctypes

import numpy as np
import pytest

from oracle import semisep

pytestmark = pytest.mark.gpu


def _pair(n, seed):
    from dwba_amd import grid as G
    g = G.exp_grid(n, 1e-5, 1e3)
    rng = np.random.default_rng(seed)
    return g.r, [rng.standard_normal(n) * g.w_simpson for _ in range(4)]


def _run(lib, hip, torch, dev, r, vecs, n_rows, L_max, noalloc):
    t = lambda x: torch.as_tensor(np.ascontiguousarray(x), dtype=torch.float64, device=dev)
    r_d = t(r)
    v_d = [t(v) for v in vecs]
    n = len(r)
    out = torch.full((2, L_max + 1), np.nan, dtype=torch.float64, device=dev)
    sh = hip.stream_handle()
    P = hip.ptr
    if noalloc:
        rc = lib.dwba_il_plain_f64_noalloc(P(r_d), n_rows, n, *[P(v) for v in v_d], L_max, P(out), sh)
        hip.check(rc, "dwba_il_plain_f64_noalloc")
    else:
        nb = lib.dwba_il_plain_f64_workspace_bytes(n_rows, n, L_max)
        ws = torch.empty(nb, dtype=torch.uint8, device=dev)
        rc = lib.dwba_il_plain_f64(P(r_d), n_rows, n, *[P(v) for v in v_d], L_max, P(out), P(ws),
                                   ctypes.c_size_t(nb), sh)
        hip.check(rc, "dwba_il_plain_f64")
    torch.cuda.synchronize()
    return out.cpu().numpy()


def test_noalloc_matches_workspace_entry_point_twice():
    import torch
    from dwba_amd import _hip
    dev = torch.device("cuda", _hip.ensure_init(0))
    lib = _hip.load_library()
    # second call smaller: fewer points, fewer rows than columns, fewer multipoles
    for (n, n_rows, L_max, seed) in ((12000, 12000, 30, 1), (2500, 1900, 12, 2)):
        r, vecs = _pair(n, seed)
        vecs[0][n_rows:] = 0.0          # alpha is zero from n_rows on
        vecs[2][n_rows:] = 0.0
        print(f"[noalloc] n={n} n_rows={n_rows} L_max={L_max}", flush=True)
        got = _run(lib, _hip, torch, dev, r, vecs, n_rows, L_max, True)
        ref = _run(lib, _hip, torch, dev, r, vecs, n_rows, L_max, False)
        assert np.array_equal(got, ref), np.max(np.abs(got - ref))
        for fam in (0, 1):
            a, b = vecs[2 * fam][:n_rows], vecs[2 * fam + 1]
            ora = semisep.il_all_L(r, np.concatenate([a, np.zeros(n - n_rows)]), b, L_max)
            rel = np.max(np.abs(got[fam] - ora) / np.abs(ora))
            assert rel <= 1e-10, (n, fam, rel)


def test_noalloc_argument_errors():
    import torch
    from dwba_amd import _hip
    dev = torch.device("cuda", _hip.ensure_init(0))
    lib = _hip.load_library()
    r, vecs = _pair(200, 3)
    t = lambda x: torch.as_tensor(x, dtype=torch.float64, device=dev)
    r_d, v_d = t(r), [t(v) for v in vecs]
    out = torch.zeros((2, 64), dtype=torch.float64, device=dev)
    P = _hip.ptr
    sh = _hip.stream_handle()
    assert lib.dwba_il_plain_f64_noalloc(P(r_d), 200, 200, *[P(v) for v in v_d], 64, P(out), sh) == _hip.DWBA_ERR_ARG
    assert b"L_max" in lib.dwba_last_error()
    assert lib.dwba_il_plain_f64_noalloc(P(r_d), 201, 200, *[P(v) for v in v_d], 4, P(out), sh) == _hip.DWBA_ERR_ARG
    assert lib.dwba_il_plain_f64_noalloc(P(r_d), 200, 200, *[P(v) for v in v_d], 4, 0, sh) == _hip.DWBA_ERR_ARG
