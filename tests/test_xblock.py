"""Block-form MFMA contraction (dwba_xblock_*).

CPU: the block decomposition the kernel evaluates (rank-one moments between
blocks of 8 points, the block's own 36 entries, FP64 scans, ranges started
from zero and joined in grid order), restated in FP64 numpy, equals the
oracle's exact contraction; the kernel's entry schedule covers the 36
(i <= j) entries of a block once each.

GPU: the FP32 and split-BF16x3 kernels against the oracle on stacked
families with ragged supports, partial groups, several ranges per group and
grids whose length is not a multiple of 8.
"""
import os
import re

import numpy as np
import pytest

from oracle import semisep

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "dwba_amd", "csrc", "xblock.hip")


def kernel_schedule():
    txt = open(SRC).read()
    a = [int(x) for x in re.search(r"constexpr int XBA\[XB_DS\] = \{([^}]*)\}", txt).group(1).split(",")]
    c = [int(x) for x in re.search(r"constexpr int XBC\[XB_DS\] = \{([^}]*)\}", txt).group(1).split(",")]
    return a, c


def test_entry_schedule_covers_block_once():
    a, c = kernel_schedule()
    seen = []
    for s in range(len(a)):
        for g in range(4):                      # lane group g reads registers rotated by g
            i, j = (a[s] + g) % 8, (c[s] + g) % 8
            seen.append((min(i, j), max(i, j)))
    assert sorted(seen) == sorted((i, j) for i in range(8) for j in range(i, 8))


def block_form(r, alpha, beta, L_max, blocks_per_range, drop_upper=False):
    """The kernel's arithmetic in FP64: per block the four moments, the
    block's own entries, the scans; ranges from zero scans, joined."""
    N = len(r)
    nb = (N + 7) // 8
    Ls = np.arange(L_max + 1, dtype=np.float64)
    rp = np.concatenate([r, np.full(8 * nb - N, np.nan)])
    ap = np.concatenate([alpha, np.zeros(8 * nb - N)])
    bp = np.concatenate([beta, np.zeros(8 * nb - N)])
    dL = np.array([semisep.diag_factor(int(L)) for L in Ls])
    n_ranges = max(1, -(-nb // blocks_per_range))
    parts = []
    for rg in range(n_ranges):
        b0, b1 = min(nb, rg * blocks_per_range), min(nb, (rg + 1) * blocks_per_range)
        S = np.zeros(L_max + 1); T = np.zeros(L_max + 1); tot = np.zeros(L_max + 1)
        Ca = np.zeros(L_max + 1); Cb = np.zeros(L_max + 1); P = np.ones(L_max + 1)
        for b in range(b0, b1):
            pts = np.arange(8 * b, 8 * b + 8)
            ok = pts < N
            rr, aa, bb = rp[pts], ap[pts], bp[pts]
            f, gl = rp[8 * b], rp[min(8 * b + 7, N - 1)]
            U = np.where(ok[None, :], (np.where(ok, rr, 1.0)[None, :] / gl) ** Ls[:, None], 0.0)
            V = np.where(ok[None, :], (f / np.where(ok, rr, 1.0)[None, :]) ** (Ls[:, None] + 1) / f, 0.0)
            mB, mA, nA, nB = U @ bb, U @ aa, V @ aa, V @ bb
            D = np.zeros(L_max + 1)
            for i in range(8):
                for j in range(i, 8):
                    if not (ok[i] and ok[j]):
                        continue
                    if i == j:
                        D += dL / rr[i] * aa[i] * bb[i]
                    else:
                        D += (rr[i] / rr[j]) ** Ls / rr[j] * (aa[i] * bb[j] + aa[j] * bb[i])
            nxt = 8 * b + 8 < N
            r1 = (f / rp[8 * b + 8]) ** Ls if nxt else np.zeros(L_max + 1)
            r2 = (gl / rp[8 * b + 8]) ** Ls if nxt else np.zeros(L_max + 1)
            tot = nA * S + (nB * T + (tot + D))
            Ca = Ca + nA * P
            Cb = Cb + nB * P
            S = r1 * S + r2 * mB
            T = r1 * T + r2 * (0.0 * mA if drop_upper else mA)
            P = P * r1
        parts.append((tot, S, T, Ca, Cb, P))
    tot = np.zeros(L_max + 1); Si = np.zeros(L_max + 1); Ti = np.zeros(L_max + 1)
    for (t0, Se, Te, Ca, Cb, P) in parts:
        tot = tot + (Ca * Si + (Cb * Ti + t0))
        Si = P * Si + Se
        Ti = P * Ti + Te
    return tot


@pytest.mark.parametrize("N,bpr", [(203, 4), (640, 1000), (997, 7)])
def test_block_form_equals_exact_contraction(N, bpr):
    rng = np.random.default_rng(N)
    r = 1e-3 * np.exp(np.linspace(0.0, np.log(2e5), N))
    a = rng.standard_normal(N) * np.exp(-r / 50)
    b = np.sin(1.3 * r + 0.4) * rng.uniform(0.5, 1.0, N)
    L_max = 20
    got = block_form(r, a, b, L_max, bpr)
    ref = semisep.il_all_L(r, a, b, L_max)
    mag = semisep.abs_bound_all_L(r, a, b, L_max)
    assert np.max(np.abs(got - ref) / mag) < 1e-13


def test_block_form_catches_a_dropped_cross_term():
    """The comparison above is sharp: the same block form without the upper
    (i < j) cross terms between blocks fails it by far."""
    N = 160
    rng = np.random.default_rng(3)
    r = np.linspace(0.05, 30.0, N)
    a = rng.standard_normal(N)
    b = rng.standard_normal(N)
    ref = semisep.il_all_L(r, a, b, 6)
    mag = semisep.abs_bound_all_L(r, a, b, 6)
    assert np.max(np.abs(block_form(r, a, b, 6, 5) - ref) / mag) < 1e-13
    assert np.max(np.abs(block_form(r, a, b, 6, 5, drop_upper=True) - ref) / mag) > 1e-2


# ---------------------------------------------------------------------------- GPU

def _families(rng, specs, L_max):
    """specs: (N, P) per grid.  Returns grids (r), stacked A, B (zero beyond
    each pair's supports) and the per-pair extents."""
    fams = []
    for (N, P) in specs:
        r = 1e-5 * np.exp(np.arange(N) * (np.log(1e8) / (N - 1)))
        ld = (N + 31) // 32 * 32
        A = np.zeros((P, ld))
        B = np.zeros((P, ld))
        ext = np.zeros(P, dtype=np.int64)
        for c in range(P):
            # supports like the exchange densities' (split index >= 3/4 of the
            # grid): integrands far below FP32's range are not representable
            rows = int(rng.integers(3 * N // 4, N + 1))
            cols = int(rng.integers(rows, N + 1))
            k1, k2 = rng.uniform(0.5, 8.0, 2)
            l1, l2 = rng.integers(0, 41, 2)
            env = (1 - np.exp(-r)) ** (l1 + 1) * (1 - np.exp(-r)) ** (l2 + 1)
            chi = np.sin(k1 * r + rng.uniform(0, np.pi)) * np.sin(k2 * r + rng.uniform(0, np.pi)) * env
            bound = r ** 2 * np.exp(-r) * rng.uniform(0.5, 2.0)
            w = np.gradient(r)
            A[c, :rows] = (chi * w)[:rows]
            B[c, :cols] = (bound * w)[:cols]
            ext[c] = cols
        fams.append((r, A, B, ext))
    return fams


@pytest.mark.gpu
@pytest.mark.parametrize("mode,bound", [("fp32", 5e-7), ("bf16x3", 5e-6)])
def test_xblock_kernel_against_oracle(mode, bound):
    import torch
    from dwba_amd import _hip, xblock
    lib = _hip.load_library()
    dev = torch.device("cuda", _hip.ensure_init())
    rng = np.random.default_rng(11)
    L_max = 30
    specs = [(1001, 70), (2400, 16), (4000, 5)]
    fams = _families(rng, specs, L_max)
    print(f"xblock {mode}: grids {specs}, L_max {L_max}", flush=True)
    T = lambda x: torch.as_tensor(np.ascontiguousarray(x), device=dev)
    r_d = [T(f[0]) for f in fams]
    tabs, grids_d = xblock.build_tables(_hip, lib, torch, dev, [(r_d[i], len(fams[i][0])) for i in range(len(fams))],
                                        mode)
    A_d = [T(f[1]) for f in fams]
    B_d = [T(f[2]) for f in fams]
    n_tot = sum(P for _, P in specs)
    nl = L_max + 1
    families, row0 = [], 0
    for i, f in enumerate(fams):
        families.append((i, A_d[i], B_d[i], f[1].shape[1], f[3], row0, 1))
        row0 += len(f[3])
    for bpr in (16, 64, 100000):
        plan = xblock.XBlockPlan(_hip, lib, torch, dev, families, tabs, mode, blocks_per_range=bpr)
        part = plan.partial(lib, torch, dev)
        out = torch.full((n_tot, 2, nl), 7.0, dtype=torch.float64, device=dev)
        plan.launch(_hip, lib, mode, L_max, part, out)
        got = out.cpu().numpy()
        assert np.all(got[:, 0, :] == 7.0)        # the other slot is untouched
        worst = np.zeros(nl)
        row = 0
        for (r, A, B, ext) in fams:
            for c in range(A.shape[0]):
                N = len(r)
                ref = semisep.il_all_L(r, A[c, :N], B[c, :N], L_max)
                mag = semisep.abs_bound_all_L(r, A[c, :N], B[c, :N], L_max)
                assert np.all(np.isfinite(got[row, 1])), (row, got[row, 1])
                worst = np.maximum(worst, np.abs(got[row, 1] - ref) / np.maximum(mag, 1e-300))
                row += 1
        print(f"  blocks/range {bpr}: max |dI| / sum|aKb| per L: {np.array2string(worst, precision=2)}",
              flush=True)
        assert np.max(worst) < bound, (bpr, worst)
