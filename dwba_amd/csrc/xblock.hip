// Stacked contraction on the matrix cores in block form:
//     I_p(L) = alpha_p^T K_L beta_p,   K_L(i, j) = r_<^L / r_>^(L+1),
// for every channel pair p of every energy of a batch and L = 0..31, with
// FP32 (v_mfma_f32_16x16x4_f32) or split-BF16x3 (v_mfma_f32_16x16x16_bf16)
// operands, FP32 accumulation inside a block of 8 grid points and FP64
// between blocks.  Same contraction as dwba_xgemm_f32 / _bf16x3 (which
// replace the per-pair exchange GEMVs of radial_ME_all_L,
// dwba_matrix_elements.py:1281-1289, and radial_ME_all_L_gpu, :2508-2509),
// without the dense tiles' redundant work.
//
// The grid is cut into blocks of 8 points.  A tile of K_L between two
// different blocks is exactly rank one (K_L(i, j) = r_j^L r_i^-(L+1) for
// j < i), so with f_b, g_b the first and last radius of block b
//     Mb_b(L) = sum_j beta_j (r_j / g_b)^L         Ma_b(L) = sum_i alpha_i (r_i / g_b)^L
//     Na_b(L) = sum_i alpha_i (f_b / r_i)^(L+1) / f_b   Nb_b(L) = sum_j beta_j (f_b / r_j)^(L+1) / f_b
//     I(L) = sum_b [ D_b(L) + Na_b(L) S_b(L) + Nb_b(L) T_b(L) ]
//     S_{b+1} = (f_b/f_{b+1})^L S_b + (g_b/f_{b+1})^L Mb_b     (T alike with Ma)
// and D_b(L) = sum_{i<=j in b} K_L(i, j) Pi_ij, Pi_ij = alpha_i beta_j +
// alpha_j beta_i (i < j), alpha_i beta_i (i = j).  Every factor is <= 1 (times
// 1/f_b).  The four moments are GEMMs (32 multipoles x 8 points) x (8 points
// x 16 pairs) and D is a GEMM (32 multipoles x 36 entries) x (36 entries x
// 16 pairs) against a per-grid table; the scans run in FP64 on the VALU.
//
// MFMA operand layout (16x16x4 f32): lane l holds A[l % 16][l / 16] and
// B[l / 16][l % 16]; D[4 (l / 16) + v][l % 16] in element v.  Rows are
// multipoles (two 16-row tiles), columns are the wavefront's 16 pairs, the k
// index of lane group g = l / 16 is one of the block's points or entries.
// Lane group g loads the block's alpha / beta rotated by g (register k holds
// point (k + g) mod 8), so every k-step reads the same registers in all
// groups: the 9 steps (a_s, c_s) below enumerate the 36 entries as orbits of
// the rotation (diagonal, distance 1, 2, 3 twice each, distance 4 once).
//
// Work: a workgroup = 4 wavefronts = 64 consecutive pairs of one stacked
// matrix over one range of blocks (the same block boundaries for every group
// of a grid); ranges start from zero scans
// and also accumulate Ca = sum_b Na_b P_b, Cb = sum_b Nb_b P_b,
// P_b = (f_start / f_b)^L, so that a join kernel adds the true incoming scans
// range by range (the scans are linear).
#include "common.hpp"

#include <type_traits>

namespace dwba {
namespace {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef short v4s __attribute__((ext_vector_type(4)));

constexpr int XBP = 8;           // grid points per block
constexpr int XB_DS = 9;         // k-steps of 4 entries (36 = 8 * 9 / 2)
constexpr int XB_NQ = 5;         // partial quantities per (pair, L): tot, S, T, Ca, Cb
constexpr int XB_GP = 64;        // pairs per workgroup
constexpr size_t XB_PSTRIDE = (size_t)XB_NQ * XB_GP * 32 + 32;   // doubles per range partial (+ P)

__constant__ int XB_A[XB_DS] = {0, 4, 0, 4, 0, 4, 0, 4, 0};
__constant__ int XB_C[XB_DS] = {0, 4, 1, 5, 2, 6, 3, 7, 4};
constexpr int XBA[XB_DS] = {0, 4, 0, 4, 0, 4, 0, 4, 0};
constexpr int XBC[XB_DS] = {0, 4, 1, 5, 2, 6, 3, 7, 4};

// Per-block table, FP32 operands.
struct XbTabF32 {
    double rho1[32];             // (f_b / f_{b+1})^L     (0 for the last block)
    double rho2[32];             // (g_b / f_{b+1})^L
    float kd[XB_DS][2][64];      // K_L of entry (s, lane / 16), L = 16 t + lane % 16
    float u[2][2][64];           // (r_k / g_b)^L,  k = 4 s + lane / 16
    float v[2][2][64];           // (f_b / r_k)^(L+1) / f_b
    float pad_[256];             // to 2 x 256 chunks of 16 bytes: every thread stages two
};
static_assert(sizeof(XbTabF32) == 8192, "table layout");

// Per-block table, split-BF16 operands for v_mfma_f32_16x16x16_bf16: lane l
// holds A[l % 16][4 (l / 16) + e], e = 0..3, as 4 bf16.  The 9 entry steps
// pack 4 per instruction (steps 4m .. 4m+3 in slot e, the last instruction
// holds step 8 in slot 0); the point moments use slots 0, 1 (points
// lane/16 and 4 + lane/16).  hi / lo: bf16(x) and bf16(x - hi).
struct XbTabBF16 {
    double rho1[32];
    double rho2[32];
    v4s kd_hi[3][2][64];
    v4s kd_lo[3][2][64];
    v4s u_hi[2][64], u_lo[2][64];
    v4s v_hi[2][64], v_lo[2][64];
    double pad_[192];            // to 3 x 256 chunks of 16 bytes
};
static_assert(sizeof(XbTabBF16) == 12288, "table layout");

__device__ __forceinline__ unsigned short bf16_bits(float x) {
    return __builtin_bit_cast(unsigned short, (__bf16)x);
}
__device__ __forceinline__ float bf16_val(unsigned short b) {
    return (float)__builtin_bit_cast(__bf16, b);
}

// K_L of entry (i, j), i <= j, of block points, FP64 (0 beyond the grid).
__device__ double xb_kernel_entry(const double* r, int N, int i, int j, int L) {
    if (i > j) { const int t = i; i = j; j = t; }
    if (j >= N) return 0.0;
    if (i == j) return (L == 0 ? 1.0 : exp((double)L * log(1.0 - 1e-12))) / r[i];
    return pow(r[i] / r[j], (double)L) / r[j];
}

// One thread per (block, tile t, lane) of one grid.
template <bool BF16>
__global__ void xblock_table_kernel(const dwba_xblock_grid* __restrict__ grids) {
    const dwba_xblock_grid Gd = grids[blockIdx.y];
    const int N = Gd.N;
    const int nb = (N + XBP - 1) / XBP;
    const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int b = (int)(gid >> 7), t = (int)((gid >> 6) & 1), l = (int)(gid & 63);
    if (b >= nb) return;
    const double* r = Gd.r;
    const int p0 = XBP * b;
    const int L = 16 * t + (l & 15), g = l >> 4;
    const double f = r[p0], gl = r[min(p0 + XBP - 1, N - 1)];
    const bool has_next = p0 + XBP < N;
    const double fn = has_next ? r[p0 + XBP] : 1.0;
    auto uval = [&](int k) { const int p = p0 + k; return p < N ? pow(r[p] / gl, (double)L) : 0.0; };
    auto vval = [&](int k) { const int p = p0 + k; return p < N ? pow(f / r[p], (double)(L + 1)) / f : 0.0; };
    auto kval = [&](int s) {
        return xb_kernel_entry(r, N, p0 + ((XB_A[s] + g) & 7), p0 + ((XB_C[s] + g) & 7), L);
    };
    if (!BF16) {
        XbTabF32* T = reinterpret_cast<XbTabF32*>(Gd.tab) + b;
        for (int s = 0; s < XB_DS; ++s) T->kd[s][t][l] = (float)kval(s);
        for (int s = 0; s < 2; ++s) {
            T->u[s][t][l] = (float)uval(4 * s + g);
            T->v[s][t][l] = (float)vval(4 * s + g);
        }
        if (t == 0 && l < 32) {
            T->rho1[l] = has_next ? pow(f / fn, (double)l) : 0.0;
            T->rho2[l] = has_next ? pow(gl / fn, (double)l) : 0.0;
        }
    } else {
        XbTabBF16* T = reinterpret_cast<XbTabBF16*>(Gd.tab) + b;
        auto split = [](double x, unsigned short& h, unsigned short& lo) {
            const float xf = (float)x;
            h = bf16_bits(xf);
            lo = bf16_bits((float)(x - (double)bf16_val(h)));
        };
        for (int m = 0; m < 3; ++m) {
            v4s hi, lo;
            for (int e = 0; e < 4; ++e) {
                const int s = 4 * m + e;
                unsigned short h = 0, q = 0;
                if (s < XB_DS) split(kval(s), h, q);
                hi[e] = (short)h;
                lo[e] = (short)q;
            }
            T->kd_hi[m][t][l] = hi;
            T->kd_lo[m][t][l] = lo;
        }
        v4s uh = {0, 0, 0, 0}, ul = {0, 0, 0, 0}, vh = {0, 0, 0, 0}, vl = {0, 0, 0, 0};
        for (int e = 0; e < 2; ++e) {
            unsigned short h, q;
            split(uval(4 * e + g), h, q);
            uh[e] = (short)h; ul[e] = (short)q;
            split(vval(4 * e + g), h, q);
            vh[e] = (short)h; vl[e] = (short)q;
        }
        T->u_hi[t][l] = uh; T->u_lo[t][l] = ul;
        T->v_hi[t][l] = vh; T->v_lo[t][l] = vl;
        if (t == 0 && l < 32) {
            T->rho1[l] = has_next ? pow(f / fn, (double)l) : 0.0;
            T->rho2[l] = has_next ? pow(gl / fn, (double)l) : 0.0;
        }
    }
}

// FP64 scan state of one lane: 8 (multipole, pair) cells, q = 4 t + v <-> L = 16 t + 4 g + v.
struct XbScan {
    double S[8], T[8], tot[8], Ca[8], Cb[8], P[8];
};

__device__ __forceinline__ void xb_scan_init(XbScan& X) {
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        X.S[q] = 0.0; X.T[q] = 0.0; X.tot[q] = 0.0; X.Ca[q] = 0.0; X.Cb[q] = 0.0; X.P[q] = 1.0;
    }
}

// One block's scan update from its moments (MFMA outputs) and D.
__device__ __forceinline__ void xb_scan_step(XbScan& X, const double* __restrict__ rho1, const double* __restrict__ rho2,
                                             int g, const v4f (&mB)[2], const v4f (&mA)[2], const v4f (&nA)[2],
                                             const v4f (&nB)[2], const v4f (&dd)[2]) {
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const double2 r1a = *reinterpret_cast<const double2*>(rho1 + 16 * t + 4 * g);
        const double2 r1b = *reinterpret_cast<const double2*>(rho1 + 16 * t + 4 * g + 2);
        const double2 r2a = *reinterpret_cast<const double2*>(rho2 + 16 * t + 4 * g);
        const double2 r2b = *reinterpret_cast<const double2*>(rho2 + 16 * t + 4 * g + 2);
        const double r1[4] = {r1a.x, r1a.y, r1b.x, r1b.y};
        const double r2[4] = {r2a.x, r2a.y, r2b.x, r2b.y};
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int q = 4 * t + v;
            const double na = (double)nA[t][v], nb = (double)nB[t][v];
            X.tot[q] = fma(na, X.S[q], fma(nb, X.T[q], X.tot[q] + (double)dd[t][v]));
            X.Ca[q] = fma(na, X.P[q], X.Ca[q]);
            X.Cb[q] = fma(nb, X.P[q], X.Cb[q]);
            X.S[q] = fma(r1[v], X.S[q], r2[v] * (double)mB[t][v]);
            X.T[q] = fma(r1[v], X.T[q], r2[v] * (double)mA[t][v]);
            X.P[q] *= r1[v];
        }
    }
}

__device__ __forceinline__ void xb_write_partial(const XbScan& X, double* __restrict__ pp, int pair, int g, bool wp) {
#pragma unroll
    for (int t = 0; t < 2; ++t) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int q = 4 * t + v, L = 16 * t + 4 * g + v;
            pp[(0 * XB_GP + pair) * 32 + L] = X.tot[q];
            pp[(1 * XB_GP + pair) * 32 + L] = X.S[q];
            pp[(2 * XB_GP + pair) * 32 + L] = X.T[q];
            pp[(3 * XB_GP + pair) * 32 + L] = X.Ca[q];
            pp[(4 * XB_GP + pair) * 32 + L] = X.Cb[q];
            if (wp) pp[(size_t)XB_NQ * XB_GP * 32 + L] = X.P[q];
        }
    }
}

// One block's inputs in LDS: its factor table and the alpha / beta of the
// workgroup's 64 pairs.  Pair p's 8 alpha and 8 beta values sit at
// ab[16 p + ((8 ab + t) ^ (2 (p & 7)))]: the XOR swizzle keeps each 16-byte
// chunk whole and spreads a wavefront's rotated reads (16 pairs x 4 lane
// groups, one point each) evenly over the LDS banks.
template <typename Tab>
struct XbBuf {
    Tab tab;
    double ab[XB_GP * 16];
};

// One block's loads in flight in registers, issued by all 256 threads of the
// workgroup as 16-byte loads: NT chunks of the table (padded to whole
// multiples of 256 chunks) and 2 of alpha / beta (64 pairs x 8 chunks).  The
// thread's source pointers are formed once (XbSrc); a block only adds its
// offset.  Pairs beyond the group load pair 0 (never used).
typedef const v4f __attribute__((address_space(1)))* gv4f;
template <typename Tab>
struct XbSrc {
    static constexpr int NCH = (int)(sizeof(Tab) / 16);
    static constexpr int NT = NCH / 256;
    static_assert(NCH % 256 == 0, "table padded to whole rounds of 256 chunks");
    gv4f tab;                    // chunk tid of block 0
    gv4f ab[2];                  // chunk of block 0 of slot tid, tid + 256
    __device__ __forceinline__ XbSrc(const Tab* t, const dwba_xblock_group& G, int tid) {
        tab = (gv4f)(reinterpret_cast<const v4f*>(t) + tid);
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int sl = tid + 256 * q;
            const int p = sl >> 3, cc = (sl & 7) ^ (p & 7);
            const int pp = p < G.n_pairs ? p : 0;
            ab[q] = (gv4f)reinterpret_cast<const v4f*>(((cc >> 2) ? G.B : G.A) + (size_t)pp * G.ld + 2 * (cc & 3));
        }
    }
};
template <typename Tab>
struct XbStage {
    static constexpr int NT = XbSrc<Tab>::NT;
    v4f t[NT], ab[2];
    __device__ __forceinline__ void load(const XbSrc<Tab>& S, int b) {
#pragma unroll
        for (int q = 0; q < NT; ++q) t[q] = S.tab[(size_t)b * XbSrc<Tab>::NCH + 256 * q];
#pragma unroll
        for (int q = 0; q < 2; ++q) ab[q] = S.ab[q][(size_t)b * (XBP / 2)];   // 8 doubles = 4 chunks per block
    }
    __device__ __forceinline__ void store(XbBuf<Tab>& buf, int tid) const {
        v4f* d4 = reinterpret_cast<v4f*>(&buf.tab);
#pragma unroll
        for (int q = 0; q < NT; ++q) d4[tid + 256 * q] = t[q];
        v4f* a4 = reinterpret_cast<v4f*>(buf.ab);
        a4[tid] = ab[0];
        a4[tid + 256] = ab[1];
    }
};

// alpha, beta of local pair p rotated by the lane group: register k <-> point (k + g) mod 8
__device__ __forceinline__ void xb_read(const double* __restrict__ ab, int p, int g, double (&a)[8], double (&c)[8]) {
    const int base = 16 * p, sw = 2 * (p & 7);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int t = (k + g) & 7;
        a[k] = ab[base + (t ^ sw)];
        c[k] = ab[base + ((8 + t) ^ sw)];
    }
}

// One block, FP32 operands.
__device__ __forceinline__ void xb_block_f32(const XbBuf<XbTabF32>& buf, int lane, int g, int p, XbScan& X) {
    double xa[8], xc[8];
    xb_read(buf.ab, p, g, xa, xc);
    float af[8], bf[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) { af[k] = (float)xa[k]; bf[k] = (float)xc[k]; }
    const XbTabF32& tb = buf.tab;
    const v4f z = {0.f, 0.f, 0.f, 0.f};
    v4f mB[2] = {z, z}, mA[2] = {z, z}, nA[2] = {z, z}, nB[2] = {z, z}, dd[2] = {z, z};
#pragma unroll
    for (int s = 0; s < 2; ++s) {
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const float uu = tb.u[s][t][lane], vv = tb.v[s][t][lane];
            mB[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(uu, bf[4 * s], mB[t], 0, 0, 0);
            mA[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(uu, af[4 * s], mA[t], 0, 0, 0);
            nA[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(vv, af[4 * s], nA[t], 0, 0, 0);
            nB[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(vv, bf[4 * s], nB[t], 0, 0, 0);
        }
    }
#pragma unroll
    for (int s = 0; s < XB_DS; ++s) {
        const int a = XBA[s], c = XBC[s];
        const float pi = a == c ? af[a] * bf[a] : fmaf(af[a], bf[c], af[c] * bf[a]);
#pragma unroll
        for (int t = 0; t < 2; ++t)
            dd[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(tb.kd[s][t][lane], pi, dd[t], 0, 0, 0);
    }
    xb_scan_step(X, tb.rho1, tb.rho2, g, mB, mA, nA, nB, dd);
}

// hi / lo bf16 of x.
__device__ __forceinline__ void xb_split(float x, short& h, short& l) {
    const unsigned short hb = bf16_bits(x);
    h = (short)hb;
    l = (short)bf16_bits(x - bf16_val(hb));
}

// One block, split-BF16x3 operands.
__device__ __forceinline__ void xb_block_bf16(const XbBuf<XbTabBF16>& buf, int lane, int g, int p, XbScan& X) {
    double xa[8], xc[8];
    xb_read(buf.ab, p, g, xa, xc);
    float af[8], bf[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) { af[k] = (float)xa[k]; bf[k] = (float)xc[k]; }
    // point operands (slots 0, 1: points g and 4 + g), split from FP64
    v4s ah = {0, 0, 0, 0}, al = {0, 0, 0, 0}, bh = {0, 0, 0, 0}, bl = {0, 0, 0, 0};
#pragma unroll
    for (int e = 0; e < 2; ++e) {
        const unsigned short ha = bf16_bits((float)xa[4 * e]), hb = bf16_bits((float)xc[4 * e]);
        ah[e] = (short)ha;
        bh[e] = (short)hb;
        al[e] = (short)bf16_bits((float)(xa[4 * e] - (double)bf16_val(ha)));
        bl[e] = (short)bf16_bits((float)(xc[4 * e] - (double)bf16_val(hb)));
    }
    const XbTabBF16& tb = buf.tab;
    const v4f z = {0.f, 0.f, 0.f, 0.f};
    v4f mB[2] = {z, z}, mA[2] = {z, z}, nA[2] = {z, z}, nB[2] = {z, z}, dd[2] = {z, z};
#define XB3(acc, A_h, A_l, B_h, B_l)                                                              \
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A_l, B_h, acc, 0, 0, 0);                       \
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A_h, B_l, acc, 0, 0, 0);                       \
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A_h, B_h, acc, 0, 0, 0);
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const v4s uh = tb.u_hi[t][lane], ul = tb.u_lo[t][lane];
        const v4s vh = tb.v_hi[t][lane], vl = tb.v_lo[t][lane];
        XB3(mB[t], uh, ul, bh, bl)
        XB3(mA[t], uh, ul, ah, al)
        XB3(nA[t], vh, vl, ah, al)
        XB3(nB[t], vh, vl, bh, bl)
    }
#pragma unroll
    for (int m = 0; m < 3; ++m) {
        v4s ph = {0, 0, 0, 0}, pl = {0, 0, 0, 0};
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int s = 4 * m + e;
            if (s < XB_DS) {
                const int a = XBA[s], c = XBC[s];
                const float pi = a == c ? af[a] * bf[a] : fmaf(af[a], bf[c], af[c] * bf[a]);
                short h, l;
                xb_split(pi, h, l);
                ph[e] = h;
                pl[e] = l;
            }
        }
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const v4s kh = tb.kd_hi[m][t][lane], kl = tb.kd_lo[m][t][lane];
            XB3(dd[t], kh, kl, ph, pl)
        }
    }
#undef XB3
    xb_scan_step(X, tb.rho1, tb.rho2, g, mB, mA, nA, nB, dd);
}

// The workgroup's 4 wavefronts (16 pairs each) sweep the same blocks of one
// grid.  Each block's table and alpha / beta go global -> registers -> LDS
// through two register stages and two LDS buffers: the loads of block b + 2
// are issued while block b is contracted and written to LDS after block
// b + 1 is, so every load has two blocks' time to arrive.  Every wavefront
// takes part in the loads and barriers, also one whose pairs are all beyond
// the group.
template <bool BF16>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void xblock_kernel(
    const dwba_xblock_group* __restrict__ groups, const int2* __restrict__ items, double* __restrict__ part) {
    using Tab = typename std::conditional<BF16, XbTabBF16, XbTabF32>::type;
    __shared__ __attribute__((aligned(16))) XbBuf<Tab> s_buf[2];
    const int2 it = items[blockIdx.x];
    if (it.x < 0) return;                          // uniform over the workgroup
    const dwba_xblock_group G = groups[it.x];
    const int tid = threadIdx.x;
    const int lane = tid & 63, w = tid >> 6;
    const int pc = lane & 15, g = lane >> 4;
    const int p = 16 * w + pc;                     // pair within the group
    const bool active = 16 * w < G.n_pairs;        // this wavefront has pairs
    // range r is blocks [r bpr, (r + 1) bpr) of the grid for every group, so the
    // groups of one grid read the same table blocks in the same ranges
    const int b0 = min(G.n_blocks, it.y * G.blocks_per_range);
    const int b1 = min(G.n_blocks, b0 + G.blocks_per_range);
    const Tab* __restrict__ tab = reinterpret_cast<const Tab*>(G.tab);
    XbScan X;
    xb_scan_init(X);
    const XbSrc<Tab> src(tab, G, tid);
    XbStage<Tab> sA, sB;
    auto block = [&](int buf) {
        if constexpr (BF16) xb_block_bf16(s_buf[buf], lane, g, p, X);
        else xb_block_f32(s_buf[buf], lane, g, p, X);
    };
    if (b0 < b1) {
        sA.load(src, b0);
        if (b0 + 1 < b1) sB.load(src, b0 + 1);
        sA.store(s_buf[0], tid);
    }
    __syncthreads();
    for (int b = b0; b < b1; b += 2) {
        // block b in buffer 0; block b + 1 in flight in sB
        if (b + 2 < b1) sA.load(src, b + 2);
        if (active) block(0);
        if (b + 1 < b1) sB.store(s_buf[1], tid);  // buffer 1 was last read before the previous barrier
        __syncthreads();
        if (b + 1 >= b1) break;
        // block b + 1 in buffer 1; block b + 2 in flight in sA
        if (b + 3 < b1) sB.load(src, b + 3);
        if (active) block(1);
        if (b + 2 < b1) sA.store(s_buf[0], tid);
        __syncthreads();
    }
    if (!active) return;
    double* __restrict__ pp = part + (size_t)(G.range0 + it.y) * XB_PSTRIDE;
    xb_write_partial(X, pp, p, g, w == 0 && pc == 0);
}

// Join of a group's ranges, in grid order: one thread per (pair, L).
__global__ void xblock_join_kernel(const dwba_xblock_group* __restrict__ groups, const double* __restrict__ part,
                                   int NL, double* __restrict__ out) {
    const dwba_xblock_group G = groups[blockIdx.x];
    const int idx = blockIdx.y * blockDim.x + threadIdx.x;
    const int pair = idx >> 5, L = idx & 31;
    if (pair >= G.n_pairs || L >= NL) return;
    double tot = 0.0, Si = 0.0, Ti = 0.0;
    for (int rg = 0; rg < G.n_ranges; ++rg) {
        const double* pp = part + (size_t)(G.range0 + rg) * XB_PSTRIDE;
        const double t0 = pp[(0 * XB_GP + pair) * 32 + L];
        const double Se = pp[(1 * XB_GP + pair) * 32 + L];
        const double Te = pp[(2 * XB_GP + pair) * 32 + L];
        const double Ca = pp[(3 * XB_GP + pair) * 32 + L];
        const double Cb = pp[(4 * XB_GP + pair) * 32 + L];
        const double P = pp[(size_t)XB_NQ * XB_GP * 32 + L];
        tot += fma(Ca, Si, fma(Cb, Ti, t0));
        Si = fma(P, Si, Se);
        Ti = fma(P, Ti, Te);
    }
    out[(size_t)(G.out_row0 + pair) * 2 * NL + (size_t)G.slot * NL + L] = tot;
}

}  // namespace
}  // namespace dwba

using namespace dwba;

extern "C" {

size_t dwba_xblock_table_bytes(int32_t N, int32_t mode) {
    if (N <= 0 || (mode != DWBA_XBLOCK_F32 && mode != DWBA_XBLOCK_BF16X3)) return 0;
    const size_t nb = (size_t)(N + XBP - 1) / XBP;
    return nb * (mode == DWBA_XBLOCK_F32 ? sizeof(XbTabF32) : sizeof(XbTabBF16));
}

size_t dwba_xblock_partial_doubles(int32_t n_items) { return n_items > 0 ? (size_t)n_items * XB_PSTRIDE : 0; }

int32_t dwba_xblock_tables(int32_t n_grids, const dwba_xblock_grid* d_grids, int32_t max_n, int32_t mode,
                           void* stream) {
    DWBA_REQUIRE(initialised(), "dwba_xblock_tables: call dwba_init first");
    DWBA_REQUIRE(n_grids > 0 && d_grids && max_n > 0, "dwba_xblock_tables: bad arguments");
    DWBA_REQUIRE(mode == DWBA_XBLOCK_F32 || mode == DWBA_XBLOCK_BF16X3, "dwba_xblock_tables: bad mode %d", mode);
    const long threads = (long)((max_n + XBP - 1) / XBP) * 128;
    const dim3 grid((unsigned)((threads + 255) / 256), (unsigned)n_grids);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (mode == DWBA_XBLOCK_F32)
        hipLaunchKernelGGL(xblock_table_kernel<false>, grid, dim3(256), 0, s, d_grids);
    else
        hipLaunchKernelGGL(xblock_table_kernel<true>, grid, dim3(256), 0, s, d_grids);
    DWBA_LAUNCH_CHECK("xblock_table_kernel");
    return DWBA_OK;
}

int32_t dwba_xblock_contract(int32_t mode, int32_t n_groups, const dwba_xblock_group* d_groups, int32_t n_items,
                             const int32_t* d_items, int32_t L_max, double* d_partial, double* d_out,
                             void* stream) {
    DWBA_REQUIRE(initialised(), "dwba_xblock_contract: call dwba_init first");
    DWBA_REQUIRE(mode == DWBA_XBLOCK_F32 || mode == DWBA_XBLOCK_BF16X3, "dwba_xblock_contract: bad mode %d", mode);
    DWBA_REQUIRE(L_max >= 0 && L_max <= 31, "dwba_xblock_contract: L_max %d outside 0..31", L_max);
    DWBA_REQUIRE(n_groups >= 0 && n_items >= 0 && (n_groups == 0 || (d_groups && d_items && d_partial && d_out)),
                 "dwba_xblock_contract: bad arguments");
    if (n_groups == 0 || n_items == 0) return DWBA_OK;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const int2* items = reinterpret_cast<const int2*>(d_items);
    if (mode == DWBA_XBLOCK_F32)
        hipLaunchKernelGGL(xblock_kernel<false>, dim3(n_items), dim3(256), 0, s, d_groups, items, d_partial);
    else
        hipLaunchKernelGGL(xblock_kernel<true>, dim3(n_items), dim3(256), 0, s, d_groups, items, d_partial);
    DWBA_LAUNCH_CHECK("xblock_kernel");
    hipLaunchKernelGGL(xblock_join_kernel, dim3(n_groups, XB_GP * 32 / 256), dim3(256), 0, s, d_groups, d_partial,
                       L_max + 1, d_out);
    DWBA_LAUNCH_CHECK("xblock_join_kernel");
    return DWBA_OK;
}

}  // extern "C"
