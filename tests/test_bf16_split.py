"""Accuracy floor of split-BF16x3 (CPU, no GPU needed).

The split-BF16x3 MFMA modes write each operand as hi + lo with hi = bf16(x),
lo = bf16(x - hi) (about 16 significand bits) and sum hi*hi + hi*lo + lo*hi.
Evaluated here in exact (FP64) arithmetic -- no accumulation error at all --
the contraction alpha^T K_L beta of such operands already differs from the
FP64 value by a few 1e-7 of sum|alpha K beta| on a sample of synthetic pairs
(the maximum grows with the sample: ~1.6e-6 over the 25,600 pairs of the
64-energy scan, as the GPU kernels measure), while FP32 operands give ~5e-9:
the error of the mode is that of its operands, not of the kernels.
"""
import numpy as np

from oracle import semisep


def bf16(x):
    """Round FP32 to bf16 (nearest, ties to even), returned as FP32."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def split(x):
    x = np.asarray(x, np.float64)
    h = bf16(x.astype(np.float32)).astype(np.float64)
    return h, bf16((x - h).astype(np.float32)).astype(np.float64)


def test_bf16_rounding_helper():
    assert bf16(np.float32(1.0)) == 1.0
    assert bf16(np.float32(1.0 + 2.0 ** -8)) == 1.0            # tie: to even
    assert bf16(np.float32(1.0 + 3 * 2.0 ** -8)) == 1.0 + 2.0 ** -6
    h, lo = split(np.pi)
    assert abs(h + lo - np.pi) <= 2.0 ** -17 * np.pi


def test_split_bf16x3_floor_is_the_operands():
    rng = np.random.default_rng(5)
    N, P, L_max = 1000, 24, 30
    r = 1e-5 * np.exp(np.arange(N) * (np.log(1e8) / (N - 1)))
    w = np.gradient(r)
    A = np.zeros((P, N))
    B = np.zeros((P, N))
    for c in range(P):
        rows = int(rng.integers(3 * N // 4, N + 1))
        l1, l2 = rng.integers(0, 41, 2)
        k1, k2 = rng.uniform(0.5, 8.0, 2)
        env = (1 - np.exp(-r)) ** (l1 + l2 + 2)
        A[c, :rows] = (np.sin(k1 * r + rng.uniform(0, np.pi)) * np.sin(k2 * r + rng.uniform(0, np.pi)) * env * w)[:rows]
        B[c] = r ** 2 * np.exp(-r) * rng.uniform(0.5, 2.0) * w
    bh, bl = split(B.T)
    b32 = B.T.astype(np.float32).astype(np.float64)
    floor3 = floor32 = 0.0
    for L in range(0, L_max + 1, 5):
        K = semisep.k_dense(r, L).astype(np.float32).astype(np.float64)   # K formed in FP32 as on the GPU
        kh, kl = split(K)
        I3 = np.einsum("pi,ip->p", A, kh @ bh + kh @ bl + kl @ bh)
        I32 = np.einsum("pi,ip->p", A, K @ b32)
        ref = np.array([semisep.il_all_L(r, A[c], B[c], L, L_min=L)[0] for c in range(P)])
        mag = np.array([semisep.il_all_L(r, np.abs(A[c]), np.abs(B[c]), L, L_min=L)[0] for c in range(P)])
        floor3 = max(floor3, np.max(np.abs(I3 - ref) / mag))
        floor32 = max(floor32, np.max(np.abs(I32 - ref) / mag))
    print(f"exact-arithmetic floor: split-BF16x3 {floor3:.2e}, FP32 operands {floor32:.2e}")
    assert floor3 > 2e-7
    assert floor32 < 2e-8
    assert floor3 > 30 * floor32
