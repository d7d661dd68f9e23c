// Stacked exchange contraction on MFMA:  I_p(L) = alpha_p^T K_L beta_p  for every
// channel pair p of every energy of a batch, one persistent launch.
//
// Per energy the beta vectors of all pairs are the columns of B (N x P); the
// kernel forms T = K_L B tile by tile on the matrix cores and contracts each
// finished column of T with its own alpha_p in FP64 (the "epilogue"), so
// neither K_L nor T is ever written to HBM.  This replaces, for a whole scan,
// the per-pair exchange GEMVs of radial_ME_all_L (dwba_matrix_elements.py:
// 1281-1289, oscillatory_integrals.py:2756) and of radial_ME_all_L_gpu
// (dwba_matrix_elements.py:2508-2509).
//
// K_L tiles come from per-row-block factor tables: with c = r at the first
// row of the block,  u_x = (r_x / c)^L,  v_x = (c / r_x)^L / r_x  (FP64 exp of
// the log-grid, rounded once to the MFMA type), and
//     K_L(i, j) = u_min(i,j) * v_max(i,j)
// which is exactly r_<^L / r_>^(L+1).  Chunks of 32 columns lie entirely left
// of, inside, or right of the 256-row block, so only the needed factor is
// formed (no overflow: u <= (r_last/r_first)^L of one block, v <= 1/r).
//
// Tiling (gfx950, wave64): workgroup = 4 waves = 256 rows x 128 pairs; a wave
// owns 64 rows x 128 pairs = 2 x 4 tiles of 32x32 (v_mfma_f32_32x32x2_f32
// for FP32 operands, v_mfma_f32_32x32x16_bf16 x3 for split-BF16).  The k
// (grid) dimension streams in chunks of 32 through double-buffered LDS.
// Work items are taken from an atomic counter in cost order (longest first).
#include "common.hpp"

namespace dwba {
namespace {

typedef float v16f __attribute__((ext_vector_type(16)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef short v8s __attribute__((ext_vector_type(8)));

constexpr int XT = 256;          // threads per workgroup
constexpr int XROWS = 256;       // rows per workgroup
constexpr int XCOLS = 128;       // pairs per workgroup
constexpr int XK = 32;           // grid points per chunk
constexpr int BPAD = 36;         // LDS row pitch (floats) of the beta tile

struct Item {
    int e, L, cb, rb;
};

__device__ inline float fac_u(double lr, double c, int L) { return (float)exp((double)L * (lr - c)); }
__device__ inline float fac_v(double lr, double c, int L) { return (float)exp((double)L * (c - lr) - lr); }

constexpr int XFOLD = 8;         // chunks (256 grid points) per FP32 accumulation run

// Two-level FP32 accumulation: the MFMA accumulators T run over XFOLD chunks
// and are then folded into a second set T2 (T2 += T, T = 0), so no FP32 sum
// is longer than 256 grid points plus the (short) sum of the folds; the FP32
// rounding error of a sum grows with its length.
__device__ __forceinline__ void fold_acc(v16f (&acc)[2][4], v16f (&acc2)[2][4]) {
#pragma unroll
    for (int rg = 0; rg < 2; ++rg)
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) {
            acc2[rg][cb] += acc[rg][cb];
#pragma unroll
            for (int q = 0; q < 16; ++q) acc[rg][cb][q] = 0.0f;
        }
}

// part[cb] = sum over this lane's rows of alpha_p[i] (T2[i, p] + T[i, p]), FP64.
__device__ __forceinline__ void contract_acc(const v16f (&acc)[2][4], const v16f (&acc2)[2][4], double (&part)[4],
                                             const dwba_xgemm_energy& E, int p0, int row0, int wave, int half,
                                             int l32, int ncb_act) {
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) {
        double sacc = 0.0;
        const int p = p0 + cb * 32 + l32;
        if (cb < ncb_act && p < E.P) {
            const double* ap = E.A + (size_t)p * E.ld;
#pragma unroll
            for (int rg = 0; rg < 2; ++rg) {
#pragma unroll
                for (int q = 0; q < 16; ++q) {
                    const int m = 8 * (q >> 2) + 4 * half + (q & 3);
                    const int i = row0 + wave * 64 + rg * 32 + m;
                    if (i < E.N) sacc = fma(ap[i], (double)acc2[rg][cb][q] + (double)acc[rg][cb][q], sacc);
                }
            }
        }
        part[cb] = sacc;
    }
}

template <int MODE>
__global__ __launch_bounds__(XT, 1) void xgemm_kernel(const dwba_xgemm_energy* __restrict__ en,
                                                      const int4* __restrict__ items, int n_items,
                                                      int* __restrict__ counter, double* __restrict__ partial) {
    __shared__ __attribute__((aligned(16))) float s_b[2][XCOLS * BPAD];
    __shared__ float s_fu[2][XK], s_fv[2][XK];
    __shared__ double s_red[4][XCOLS];
    __shared__ int s_item;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int half = lane >> 5, l32 = lane & 31;

    for (;;) {
        __syncthreads();
        if (tid == 0) s_item = atomicAdd(counter, 1);
        __syncthreads();
        const int it = s_item;
        if (it >= n_items) break;
        const int4 w4 = items[it];
        const Item I{w4.x, w4.y, w4.z, w4.w};
        const dwba_xgemm_energy E = en[I.e];
        const int L = I.L;
        const int row0 = I.rb * XROWS;
        const int p0 = I.cb * XCOLS;
        const int jmax = E.col_extent[I.cb];
        const int nchunk = (jmax + XK - 1) / XK;
        const int ncb_act = (min(XCOLS, E.P - p0) + 31) / 32;   // active 32-wide column tiles
        const double c0 = E.logr[row0];

        // this lane's rows and their factors
        float ru[2], rv[2];
        int rowi[2];
#pragma unroll
        for (int rg = 0; rg < 2; ++rg) {
            const int i = row0 + wave * 64 + rg * 32 + l32;
            rowi[rg] = i;
            if (i < E.N) {
                const double lr = E.logr[i];
                ru[rg] = fac_u(lr, c0, L);
                rv[rg] = fac_v(lr, c0, L);
            } else {
                ru[rg] = 0.0f;
                rv[rg] = 0.0f;
            }
        }

        v16f acc[2][4];
#pragma unroll
        for (int rg = 0; rg < 2; ++rg)
#pragma unroll
            for (int cb = 0; cb < 4; ++cb)
#pragma unroll
                for (int q = 0; q < 16; ++q) acc[rg][cb][q] = 0.0f;
        v16f acc2[2][4];
#pragma unroll
        for (int rg = 0; rg < 2; ++rg)
#pragma unroll
            for (int cb = 0; cb < 4; ++cb)
#pragma unroll
                for (int q = 0; q < 16; ++q) acc2[rg][cb][q] = 0.0f;
        double part[4];

        // ---- staging of chunk c: beta tile (128 pairs x 32 points) + column factors
        v4f stage[4];
        auto load_chunk = [&](int c) {
            const int jc = c * XK;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int idx = tid + XT * q;        // 0..1023 float4s
                const int col = idx >> 3, part = idx & 7;
                const int p = p0 + col;
                if (p < E.P)
                    stage[q] = *reinterpret_cast<const v4f*>(E.B + (size_t)p * E.ld + jc + 4 * part);
                else
                    stage[q] = v4f{0.f, 0.f, 0.f, 0.f};
            }
        };
        auto store_chunk = [&](int c, int buf) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int idx = tid + XT * q;
                const int col = idx >> 3, part = idx & 7;
                *reinterpret_cast<v4f*>(&s_b[buf][col * BPAD + 4 * part]) = stage[q];
            }
            if (tid < 2 * XK) {   // wave 0 forms the column factors
                const int jj = tid & (XK - 1);
                const int j = c * XK + jj;
                const bool is_u = tid < XK;
                float f = 0.0f;
                if (j < E.N) {
                    const double lr = E.logr[j];
                    const bool left = j < row0 + XROWS;    // u needed for j < i
                    const bool right = j >= row0;          // v needed for j > i
                    if (is_u && left) f = fac_u(lr, c0, L);
                    if (!is_u && right) f = fac_v(lr, c0, L);
                }
                if (is_u) s_fu[buf][jj] = f; else s_fv[buf][jj] = f;
            }
        };

        if (nchunk > 0) {
            load_chunk(0);
            store_chunk(0, 0);
        }
        __syncthreads();
        for (int c = 0; c < nchunk; ++c) {
            const int buf = c & 1;
            if (c + 1 < nchunk) load_chunk(c + 1);
            const int jc = c * XK;
            const bool left = jc + XK <= row0;              // all j < all rows of the block
            const bool right = jc >= row0 + XROWS;          // all j > all rows
#pragma unroll
            for (int s4 = 0; s4 < 16; s4 += 4) {
                v4f bq[4];
#pragma unroll
                for (int cb = 0; cb < 4; ++cb)
                    bq[cb] = *reinterpret_cast<const v4f*>(&s_b[buf][(cb * 32 + l32) * BPAD + 16 * half + s4]);
#pragma unroll
                for (int ss = 0; ss < 4; ++ss) {
                    const int kk = 16 * half + s4 + ss;   // column of this lane in the chunk
                    const int j = jc + kk;
                    const float fu = s_fu[buf][kk], fv = s_fv[buf][kk];
                    float a[2];
#pragma unroll
                    for (int rg = 0; rg < 2; ++rg) {
                        float kv;
                        if (right) kv = ru[rg] * fv;
                        else if (left) kv = fu * rv[rg];
                        else {
                            const int i = rowi[rg];
                            kv = j > i ? ru[rg] * fv : (j < i ? fu * rv[rg] : ru[rg] * rv[rg]);
                        }
                        a[rg] = kv;
                    }
#pragma unroll
                    for (int cb = 0; cb < 4; ++cb) {
                        if (cb < ncb_act) {
#pragma unroll
                            for (int rg = 0; rg < 2; ++rg)
                                acc[rg][cb] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[rg], bq[cb][ss], acc[rg][cb], 0, 0, 0);
                        }
                    }
                }
            }
            if ((c & (XFOLD - 1)) == XFOLD - 1 && c + 1 < nchunk) fold_acc(acc, acc2);
            // buf^1 was last read in iteration c-1, which ended with a barrier
            if (c + 1 < nchunk) store_chunk(c + 1, buf ^ 1);
            __syncthreads();
        }

        // ---- epilogue: partial_p = sum over this block's rows of alpha_p[i] T[i, p]  (FP64)
        contract_acc(acc, acc2, part, E, p0, row0, wave, half, l32, ncb_act);
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) part[cb] += __shfl_xor(part[cb], 32, 64);
        if (half == 0) {
#pragma unroll
            for (int cb = 0; cb < 4; ++cb) s_red[wave][cb * 32 + l32] = part[cb];
        }
        __syncthreads();
        if (tid < XCOLS) {
            const double v = ((s_red[0][tid] + s_red[1][tid]) + s_red[2][tid]) + s_red[3][tid];
            const size_t off = E.part_off + ((size_t)(I.L * E.n_cb + I.cb) * E.n_rb + I.rb) * XCOLS;
            partial[off + tid] = v;
        }
    }
}

constexpr int HPITCH = 40;       // LDS row pitch (bf16 elements) of the split-beta tiles

// Split-BF16x3 variant: K and beta are each split into bf16 hi + lo parts and
// T += Khi*Bhi + Khi*Blo + Klo*Bhi on v_mfma_f32_32x32x16_bf16 (FP32
// accumulation).  Same tiling, work list and epilogue as the FP32 kernel.
__global__ __launch_bounds__(XT, 1) void xgemm_bf16x3_kernel(const dwba_xgemm_energy* __restrict__ en,
                                                             const int4* __restrict__ items, int n_items,
                                                             int* __restrict__ counter,
                                                             double* __restrict__ partial) {
    __shared__ __attribute__((aligned(16))) unsigned short s_bh[2][XCOLS * HPITCH];
    __shared__ __attribute__((aligned(16))) unsigned short s_bl[2][XCOLS * HPITCH];
    __shared__ __attribute__((aligned(16))) float s_fu[2][XK], s_fv[2][XK];
    __shared__ double s_red[4][XCOLS];
    __shared__ int s_item;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int half = lane >> 5, l32 = lane & 31;

    for (;;) {
        __syncthreads();
        if (tid == 0) s_item = atomicAdd(counter, 1);
        __syncthreads();
        const int it = s_item;
        if (it >= n_items) break;
        const int4 w4 = items[it];
        const Item I{w4.x, w4.y, w4.z, w4.w};
        const dwba_xgemm_energy E = en[I.e];
        const int L = I.L;
        const int row0 = I.rb * XROWS;
        const int p0 = I.cb * XCOLS;
        const int jmax = E.col_extent[I.cb];
        const int nchunk = (jmax + XK - 1) / XK;
        const int ncb_act = (min(XCOLS, E.P - p0) + 31) / 32;
        const double c0 = E.logr[row0];

        float ru[2], rv[2];
        int rowi[2];
#pragma unroll
        for (int rg = 0; rg < 2; ++rg) {
            const int i = row0 + wave * 64 + rg * 32 + l32;
            rowi[rg] = i;
            if (i < E.N) {
                const double lr = E.logr[i];
                ru[rg] = fac_u(lr, c0, L);
                rv[rg] = fac_v(lr, c0, L);
            } else {
                ru[rg] = 0.0f;
                rv[rg] = 0.0f;
            }
        }

        v16f acc[2][4];
#pragma unroll
        for (int rg = 0; rg < 2; ++rg)
#pragma unroll
            for (int cb = 0; cb < 4; ++cb)
#pragma unroll
                for (int q = 0; q < 16; ++q) acc[rg][cb][q] = 0.0f;
        v16f acc2[2][4];
#pragma unroll
        for (int rg = 0; rg < 2; ++rg)
#pragma unroll
            for (int cb = 0; cb < 4; ++cb)
#pragma unroll
                for (int q = 0; q < 16; ++q) acc2[rg][cb][q] = 0.0f;
        double part[4];

        // chunk staging: 128 pairs x 32 points of hi and lo (64 B each per pair)
        v4f stage[4];
        auto load_chunk = [&](int c) {
            const int jc = c * XK;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int idx = tid + XT * q;          // 0..1023: [hi|lo][col][part of 4]
                const int lohi = idx >> 9;
                const int col = (idx >> 2) & 127, part = idx & 3;
                const int p = p0 + col;
                const unsigned short* src = lohi ? E.Bl : E.Bh;
                if (p < E.P)
                    stage[q] = *reinterpret_cast<const v4f*>(src + (size_t)p * E.ld + jc + 8 * part);
                else
                    stage[q] = v4f{0.f, 0.f, 0.f, 0.f};
            }
        };
        auto store_chunk = [&](int c, int buf) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int idx = tid + XT * q;
                const int lohi = idx >> 9;
                const int col = (idx >> 2) & 127, part = idx & 3;
                unsigned short* dst = lohi ? s_bl[buf] : s_bh[buf];
                *reinterpret_cast<v4f*>(dst + col * HPITCH + 8 * part) = stage[q];
            }
            if (tid < 2 * XK) {
                const int jj = tid & (XK - 1);
                const int j = c * XK + jj;
                const bool is_u = tid < XK;
                float f = 0.0f;
                if (j < E.N) {
                    const double lr = E.logr[j];
                    if (is_u && j < row0 + XROWS) f = fac_u(lr, c0, L);
                    if (!is_u && j >= row0) f = fac_v(lr, c0, L);
                }
                if (is_u) s_fu[buf][jj] = f; else s_fv[buf][jj] = f;
            }
        };

        if (nchunk > 0) {
            load_chunk(0);
            store_chunk(0, 0);
        }
        __syncthreads();
        for (int c = 0; c < nchunk; ++c) {
            const int buf = c & 1;
            if (c + 1 < nchunk) load_chunk(c + 1);
            const int jc = c * XK;
            const bool left = jc + XK <= row0;
            const bool right = jc >= row0 + XROWS;
#pragma unroll
            for (int st = 0; st < 2; ++st) {            // two k-steps of 16
                const int kb = 16 * st + 8 * half;      // this lane's 8 columns in the chunk
                const v4f fu0 = *reinterpret_cast<const v4f*>(&s_fu[buf][kb]);
                const v4f fu1 = *reinterpret_cast<const v4f*>(&s_fu[buf][kb + 4]);
                const v4f fv0 = *reinterpret_cast<const v4f*>(&s_fv[buf][kb]);
                const v4f fv1 = *reinterpret_cast<const v4f*>(&s_fv[buf][kb + 4]);
                v8bf ah[2], al[2];
#pragma unroll
                for (int rg = 0; rg < 2; ++rg) {
#pragma unroll
                    for (int t = 0; t < 8; ++t) {
                        const float fu = t < 4 ? fu0[t] : fu1[t - 4];
                        const float fv = t < 4 ? fv0[t] : fv1[t - 4];
                        float kv;
                        if (right) kv = ru[rg] * fv;
                        else if (left) kv = fu * rv[rg];
                        else {
                            const int j = jc + kb + t, i = rowi[rg];
                            kv = j > i ? ru[rg] * fv : (j < i ? fu * rv[rg] : ru[rg] * rv[rg]);
                        }
                        const __bf16 h = (__bf16)kv;
                        ah[rg][t] = h;
                        al[rg][t] = (__bf16)(kv - (float)h);
                    }
                }
#pragma unroll
                for (int cb = 0; cb < 4; ++cb) {
                    if (cb < ncb_act) {
                        const int off = (cb * 32 + l32) * HPITCH + kb;
                        const v8bf bh = *reinterpret_cast<const v8bf*>(&s_bh[buf][off]);
                        const v8bf bl = *reinterpret_cast<const v8bf*>(&s_bl[buf][off]);
#pragma unroll
                        for (int rg = 0; rg < 2; ++rg) {
                            acc[rg][cb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[rg], bh, acc[rg][cb], 0, 0, 0);
                            acc[rg][cb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[rg], bl, acc[rg][cb], 0, 0, 0);
                            acc[rg][cb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[rg], bh, acc[rg][cb], 0, 0, 0);
                        }
                    }
                }
            }
            if ((c & (XFOLD - 1)) == XFOLD - 1 && c + 1 < nchunk) fold_acc(acc, acc2);
            if (c + 1 < nchunk) store_chunk(c + 1, buf ^ 1);
            __syncthreads();
        }

        contract_acc(acc, acc2, part, E, p0, row0, wave, half, l32, ncb_act);
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) part[cb] += __shfl_xor(part[cb], 32, 64);
        if (half == 0) {
#pragma unroll
            for (int cb = 0; cb < 4; ++cb) s_red[wave][cb * 32 + l32] = part[cb];
        }
        __syncthreads();
        if (tid < XCOLS) {
            const double v = ((s_red[0][tid] + s_red[1][tid]) + s_red[2][tid]) + s_red[3][tid];
            const size_t off = E.part_off + ((size_t)(I.L * E.n_cb + I.cb) * E.n_rb + I.rb) * XCOLS;
            partial[off + tid] = v;
        }
    }
}

// hi = bf16(x), lo = bf16(x - hi) (round to nearest even), from FP64.
__global__ void split_bf16_kernel(const double* __restrict__ x, unsigned short* __restrict__ hi,
                                  unsigned short* __restrict__ lo, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double v = x[i];
        const __bf16 h = (__bf16)(float)v;
        const __bf16 l = (__bf16)(float)(v - (double)(float)h);
        hi[i] = __builtin_bit_cast(unsigned short, h);
        lo[i] = __builtin_bit_cast(unsigned short, l);
    }
}

// Sum the row-block partials of (energy, L, column block) into out[pair][1][L].
__global__ void xgemm_reduce(const dwba_xgemm_energy* __restrict__ en, const double* __restrict__ partial,
                             int L_max, double* __restrict__ out) {
    const int e = blockIdx.x, L = blockIdx.y, cb = blockIdx.z;
    const dwba_xgemm_energy E = en[e];
    if (cb >= E.n_cb) return;
    const int t = threadIdx.x;
    const int p = cb * XCOLS + t;
    if (p >= E.P) return;
    const int nrb = E.row_blocks[cb];
    double s = 0.0;
    const size_t base = E.part_off + ((size_t)(L * E.n_cb + cb) * E.n_rb) * XCOLS;
    for (int rb = 0; rb < nrb; ++rb) s += partial[base + (size_t)rb * XCOLS + t];
    out[((size_t)E.pair_base + p) * 2 * (L_max + 1) + (size_t)E.out_slot * (L_max + 1) + L] = s;
}

__global__ void fill_int(int* p, int v) { *p = v; }

__global__ void cast_f32_kernel(const double* __restrict__ x, float* __restrict__ y, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = (float)x[i];
}

}  // namespace
}  // namespace dwba

using namespace dwba;

extern "C" {

int32_t dwba_cast_f64_f32(const double* d_x, float* d_y, int64_t n, void* stream) {
    DWBA_REQUIRE(initialised(), "dwba_cast_f64_f32: call dwba_init first");
    DWBA_REQUIRE(n >= 0 && (n == 0 || (d_x && d_y)), "dwba_cast_f64_f32: bad arguments");
    if (n == 0) return DWBA_OK;
    const int64_t nb = std::min<int64_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(cast_f32_kernel, dim3((unsigned)nb), dim3(256), 0, reinterpret_cast<hipStream_t>(stream),
                       d_x, d_y, n);
    DWBA_LAUNCH_CHECK("cast_f32_kernel");
    return DWBA_OK;
}

int32_t dwba_split_bf16(const double* d_x, uint16_t* d_hi, uint16_t* d_lo, int64_t n, void* stream) {
    DWBA_REQUIRE(initialised(), "dwba_split_bf16: call dwba_init first");
    DWBA_REQUIRE(n >= 0 && (n == 0 || (d_x && d_hi && d_lo)), "dwba_split_bf16: bad arguments");
    if (n == 0) return DWBA_OK;
    const int64_t nb = std::min<int64_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(split_bf16_kernel, dim3((unsigned)nb), dim3(256), 0, reinterpret_cast<hipStream_t>(stream),
                       d_x, reinterpret_cast<unsigned short*>(d_hi), reinterpret_cast<unsigned short*>(d_lo), n);
    DWBA_LAUNCH_CHECK("split_bf16_kernel");
    return DWBA_OK;
}

size_t dwba_xgemm_partial_doubles(int32_t n_rb, int32_t n_cb, int32_t L_max) {
    return (size_t)(L_max + 1) * n_cb * n_rb * XCOLS;
}

static int32_t xgemm_launch(int mode, int32_t n_energies, const dwba_xgemm_energy* d_energies, int32_t n_items,
                            const int32_t* d_items, int32_t L_max, double* d_partial, int32_t* d_counter,
                            int32_t max_cb, double* d_out, int32_t n_workgroups, void* stream) {
    DWBA_REQUIRE(initialised(), "dwba_xgemm: call dwba_init first");
    DWBA_REQUIRE(n_energies > 0 && d_energies && d_partial && d_counter && d_out && L_max >= 0 && max_cb > 0,
                 "dwba_xgemm_f32: bad arguments");
    DWBA_REQUIRE(n_items >= 0 && (n_items == 0 || d_items), "dwba_xgemm_f32: bad work list");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(fill_int, dim3(1), dim3(1), 0, s, d_counter, 0);
    DWBA_LAUNCH_CHECK("fill_int");
    if (n_items > 0) {
        int nwg = n_workgroups > 0 ? n_workgroups : device_cu_count();
        if (nwg > n_items) nwg = n_items;
        if (mode == 0)
            hipLaunchKernelGGL(xgemm_kernel<0>, dim3(nwg), dim3(XT), 0, s, d_energies,
                               reinterpret_cast<const int4*>(d_items), n_items, d_counter, d_partial);
        else
            hipLaunchKernelGGL(xgemm_bf16x3_kernel, dim3(nwg), dim3(XT), 0, s, d_energies,
                               reinterpret_cast<const int4*>(d_items), n_items, d_counter, d_partial);
        DWBA_LAUNCH_CHECK("xgemm_kernel");
    }
    hipLaunchKernelGGL(xgemm_reduce, dim3(n_energies, L_max + 1, max_cb), dim3(XCOLS), 0, s, d_energies,
                       d_partial, L_max, d_out);
    DWBA_LAUNCH_CHECK("xgemm_reduce");
    return DWBA_OK;
}

int32_t dwba_xgemm_f32(int32_t n_energies, const dwba_xgemm_energy* d_energies, int32_t n_items,
                       const int32_t* d_items, int32_t L_max, double* d_partial, int32_t* d_counter,
                       int32_t max_cb, double* d_out, int32_t n_workgroups, void* stream) {
    return xgemm_launch(0, n_energies, d_energies, n_items, d_items, L_max, d_partial, d_counter, max_cb, d_out,
                        n_workgroups, stream);
}

int32_t dwba_xgemm_bf16x3(int32_t n_energies, const dwba_xgemm_energy* d_energies, int32_t n_items,
                          const int32_t* d_items, int32_t L_max, double* d_partial, int32_t* d_counter,
                          int32_t max_cb, double* d_out, int32_t n_workgroups, void* stream) {
    return xgemm_launch(1, n_energies, d_energies, n_items, d_items, L_max, d_partial, d_counter, max_cb, d_out,
                        n_workgroups, stream);
}

}  // extern "C"
