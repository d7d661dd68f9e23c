// Direct family of a batch, FP64:  t_L = K_L beta_dir  (one shared beta per
// energy: u_f u_i w on the full grid, dwba_matrix_elements.py:1136-1160,
// oscillatory_integrals.py:2571), then I_p(L) = alpha_dir,p^T t_L for every
// pair of the energy.  The work formula W = 2 N1 N2 min(D1, D2) + 2 N1 P with
// D2 = 1 makes this a matrix-vector product, so it runs on the FP64 VALUs in
// every precision mode; only the stacked exchange GEMM uses the matrix cores.
//
// Also the per-pair assembly of a batch (dwba_batch_finalize): tails and
// L = 0 corrections added in the order of radial_ME_all_L
// (dwba_matrix_elements.py:1112-1298).
#include "common.hpp"

namespace dwba {
namespace {

constexpr double KCLAMP = 1.0 - 1e-12;
constexpr int GT = 256;       // rows per workgroup (one per thread)
constexpr int GCH = 256;      // columns staged per LDS round

template <int NL>
__global__ __launch_bounds__(GT, 2) void dgemv_kernel(const dwba_dgemv_energy* __restrict__ en) {
    __shared__ double s_r[GCH], s_ir[GCH], s_b[GCH];
    const dwba_dgemv_energy E = en[blockIdx.y];
    const int row0 = blockIdx.x * GT;
    if (row0 >= E.rows) return;
    const int i = row0 + threadIdx.x;
    const bool valid = i < E.rows;
    const double ri = valid ? E.r[i] : 1.0;
    const double iri = 1.0 / ri;
    double acc[NL];
#pragma unroll
    for (int L = 0; L < NL; ++L) acc[L] = 0.0;
    for (int j0 = 0; j0 < E.N; j0 += GCH) {
        __syncthreads();
        {
            const int j = j0 + threadIdx.x;
            const bool in = j < E.N;
            const double rj = in ? E.r[j] : 1.0;
            s_r[threadIdx.x] = rj;
            s_ir[threadIdx.x] = 1.0 / rj;
            s_b[threadIdx.x] = in ? E.beta[j] : 0.0;
        }
        __syncthreads();
        const int jn = min(GCH, E.N - j0);
#pragma unroll 1
        for (int c = 0; c < jn; ++c) {
            const int j = j0 + c;
            const double b = s_b[c];
            double K, q;
            if (j > i) { K = s_ir[c]; q = fmin(ri * s_ir[c], KCLAMP); }
            else if (j < i) { K = iri; q = fmin(s_r[c] * iri, KCLAMP); }
            else { K = iri; q = KCLAMP; }
#pragma unroll
            for (int L = 0; L < NL; ++L) {
                acc[L] = fma(b, K, acc[L]);
                asm volatile("" : "+v"(K), "+v"(acc[L]));
                K *= q;
            }
        }
    }
    if (valid) {
#pragma unroll
        for (int L = 0; L < NL; ++L)
            if (L <= E.L_max) E.t[(size_t)L * E.ld + i] = acc[L];
    }
}

// I_p(L) = sum_i alpha_p[i] t_L[i], one workgroup per pair.
template <int NL>
__global__ __launch_bounds__(256) void ddot_kernel(const dwba_ddot_pair* __restrict__ pairs, int L_max,
                                                   double* __restrict__ out) {
    __shared__ double s_red[4][NL];
    const dwba_ddot_pair P = pairs[blockIdx.x];
    double acc[NL];
#pragma unroll
    for (int L = 0; L < NL; ++L) acc[L] = 0.0;
    for (int i = threadIdx.x; i < P.n; i += 256) {
        const double a = P.alpha[i];
        if (a != 0.0) {
#pragma unroll
            for (int L = 0; L < NL; ++L)
                if (L <= L_max) acc[L] = fma(a, P.t[(size_t)L * P.ld + i], acc[L]);
        }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int L = 0; L < NL; ++L) {
        const double v = wave_sum(acc[L]);
        if (lane == 0) s_red[wave][L] = v;
    }
    __syncthreads();
    for (int L = threadIdx.x; L <= L_max; L += 256)
        out[(size_t)P.out_index * 2 * (L_max + 1) + L] = ((s_red[0][L] + s_red[1][L]) + s_red[2][L]) + s_red[3][L];
}

// Per-pair assembly (see dwba_batch_finalize in the header).
__global__ void finalize_kernel(const dwba_finalize_pair* __restrict__ pairs, int n_pairs, int L_max,
                                double* __restrict__ out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_pairs) return;
    const dwba_finalize_pair P = pairs[k];
    const int nl = L_max + 1;
    double* o = out + (size_t)P.out_index * 2 * nl;
    const double* sc = P.scalars;
    for (int L = 0; L <= L_max; ++L) {
        double d = o[L], e = o[nl + L];
        if (P.oscillatory) {
            if (!isfinite(d)) d = 0.0;
            if (!isfinite(e)) e = 0.0;
            if (P.tail_kind == DWBA_TAIL_OUTER) d = d + P.tails[L];
        }
        if (L == 0) {
            if (fabs(sc[0]) > 1e-12) d += sc[1] * sc[0];
            if (fabs(sc[2]) > 1e-12) e += sc[3] * sc[2];
        }
        if (P.tail_kind == DWBA_TAIL_LEGACY && L >= 1) d += P.tails[L];
        o[L] = d;
        o[nl + L] = e;
    }
}

// Tiled form of ddot for pairs that share t (one energy): a workgroup takes
// up to 64 pairs of one energy; per 64-point chunk the 64 alpha rows and the
// 32 t rows are staged in LDS (coalesced loads, each read from HBM/L2 once per
// tile instead of once per pair), and thread (pair, L-group of 8) accumulates
// 8 multipoles; the next chunk's loads are in flight while the current one is
// summed.  items[k] = {first pair, pair count}; the pairs of an item have
// the same t, ld and L_max <= 31.
constexpr int TP = 64;     // pairs per tile
constexpr int TK = 64;     // grid points per chunk

__global__ __launch_bounds__(256) void ddot_tiled_kernel(const dwba_ddot_pair* __restrict__ pairs,
                                                         const int2* __restrict__ items, int L_max,
                                                         double* __restrict__ out) {
    __shared__ double s_a[TK][TP + 1];
    __shared__ __attribute__((aligned(16))) double s_t[TK][32];   // [point][L]: 8 L of a group in 4 b128 reads
    __shared__ const double* s_ap[TP];
    __shared__ int s_n;
    const int2 it = items[blockIdx.x];
    const int first = it.x, count = it.y;
    const int tid = threadIdx.x;
    const int p = tid & 63, g = tid >> 6;          // pair in the tile, L-group (a wavefront)
    const double* t = pairs[first].t;
    const int ld = pairs[first].ld;
    if (tid < TP) s_ap[tid] = tid < count ? pairs[first + tid].alpha : nullptr;
    if (tid == 0) {
        int m = 0;
        for (int q = 0; q < count; ++q) m = max(m, pairs[first + q].n);
        s_n = m;
    }
    __syncthreads();
    const int nmax = s_n;
    const int nl = L_max + 1;
    double acc[8];
#pragma unroll
    for (int l = 0; l < 8; ++l) acc[l] = 0.0;
    // chunk i0 is staged in registers while chunk i0 - TK is summed from LDS
    constexpr int QA = TP * TK / 256, QT = 32 * TK / 256;
    double ra[QA], rt[QT];
    auto load = [&](int i0) {
#pragma unroll
        for (int q = 0; q < QA; ++q) {
            const int idx = tid + 256 * q;
            const int pr = idx >> 6, pt = idx & 63;
            const int i = i0 + pt;
            ra[q] = (pr < count && i < nmax) ? s_ap[pr][i] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < QT; ++q) {
            const int idx = tid + 256 * q;
            const int L = idx >> 6, pt = idx & 63;
            const int i = i0 + pt;
            rt[q] = (L < nl && i < nmax) ? t[(size_t)L * ld + i] : 0.0;
        }
    };
    // alpha is zero beyond each pair's n (and up to ld), so nmax bounds every row
    if (nmax > 0) load(0);
    for (int i0 = 0; i0 < nmax; i0 += TK) {
#pragma unroll
        for (int q = 0; q < QA; ++q) {
            const int idx = tid + 256 * q;
            s_a[idx & 63][idx >> 6] = ra[q];
        }
#pragma unroll
        for (int q = 0; q < QT; ++q) {
            const int idx = tid + 256 * q;
            s_t[idx & 63][idx >> 6] = rt[q];
        }
        __syncthreads();
        if (i0 + TK < nmax) load(i0 + TK);
#pragma unroll 4
        for (int pt = 0; pt < TK; ++pt) {
            const double a = s_a[pt][p];
            const double2* tv = reinterpret_cast<const double2*>(&s_t[pt][8 * g]);
#pragma unroll
            for (int h = 0; h < 4; ++h) {
                const double2 v = tv[h];
                acc[2 * h] = fma(a, v.x, acc[2 * h]);
                acc[2 * h + 1] = fma(a, v.y, acc[2 * h + 1]);
            }
        }
        __syncthreads();
    }
    if (p < count) {
        double* o = out + (size_t)pairs[first + p].out_index * 2 * nl;
#pragma unroll
        for (int l = 0; l < 8; ++l) {
            const int L = 8 * g + l;
            if (L < nl) o[L] = acc[l];
        }
    }
}

// The same tiled dots on the FP64 matrix cores: for an item of <= 16 pairs
// sharing t, out[p][L] = sum_i alpha_p[i] t[L][i] is a 16 x 32 x n GEMM.
// v_mfma_f64_16x16x4_f64 takes A (16 x 4) with A[m][k] in lane m + 16k and
// B (4 x 16) with B[k][j] in lane j + 16k, and returns D (16 x 16) with
// D[lane/16 + 4v][lane % 16] in element v (measured on gfx950; the FP32
// 16x16x4 form interleaves rows the other way).  The order of the summed
// points is free as long as A and B agree, so lane (m, k) reads MS
// consecutive points k*MS .. k*MS+MS-1 of a 4*MS-point step straight from
// global memory (pair m for A, multipoles m and m + 16 for B): no LDS, no
// broadcast reads.  The MW wavefronts of a workgroup take interleaved steps
// and are summed through LDS at the end.
typedef double v4d __attribute__((ext_vector_type(4)));
constexpr int MQ = 16;        // pairs per item
constexpr int MS = 4;         // points per lane and step
constexpr int MSTEP = 4 * MS; // points per wavefront step
constexpr int MW = 4;         // wavefronts per workgroup

// NT = 1 or 2 tiles of 16 multipoles, fixed at compile time: no branch may
// separate MFMAs from the reads of their accumulators (the wait states the
// compiler inserts before those reads are only right on straight-line code).
template <int NT>
__global__ __launch_bounds__(64 * MW) void ddot_mfma_kernel(const dwba_ddot_pair* __restrict__ pairs,
                                                            const int2* __restrict__ items, int n_items, int L_max,
                                                            double* __restrict__ out) {
    __shared__ double s_red[MW - 1][4 * NT][64];
    // consecutive items (one energy's pairs, sharing t) on one XCD: blocks
    // are dealt to the 8 XCDs round-robin
    const int nb = gridDim.x, b = blockIdx.x;
    const int per = nb / 8, rem = nb % 8, xcd = b % 8;
    const int item = (xcd < rem ? xcd * (per + 1) : rem * (per + 1) + (xcd - rem) * per) + b / 8;
    if (item >= n_items) return;
    const int2 it = items[item];
    const int first = it.x, count = it.y;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int m = lane & 15, kq = lane >> 4;
    const int nl = L_max + 1;
    const double* t = pairs[first].t;
    const size_t ld = (size_t)pairs[first].ld;
    int nmax = 0;
    for (int q = 0; q < count; ++q) nmax = max(nmax, pairs[first + q].n);
    // lanes past the item's pairs or past L_max read a valid row (pair 0,
    // multipole 0) whose results land in output rows / columns never stored
    const double* ap = pairs[first + (m < count ? m : 0)].alpha;
    const double* tp[NT];
#pragma unroll
    for (int c = 0; c < NT; ++c) tp[c] = t + (size_t)(m + 16 * c < nl ? m + 16 * c : 0) * ld;
    v4d acc[NT];
#pragma unroll
    for (int c = 0; c < NT; ++c) acc[c] = v4d{0.0, 0.0, 0.0, 0.0};
    const int n_full = nmax / MSTEP;
    // one step: MS points per lane, loaded then multiplied, branch-free
    auto step_full = [&](int st) {
        const int i = st * MSTEP + kq * MS;
        double a[MS], bb[NT][MS];
#pragma unroll
        for (int u = 0; u < MS; ++u) {
            a[u] = ap[i + u];
#pragma unroll
            for (int c = 0; c < NT; ++c) bb[c][u] = tp[c][i + u];
        }
#pragma unroll
        for (int u = 0; u < MS; ++u)
#pragma unroll
            for (int c = 0; c < NT; ++c) acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u], bb[c][u], acc[c], 0, 0, 0);
    };
    int st = w;
    for (; st + MW < n_full; st += 2 * MW) {
        step_full(st);
        step_full(st + MW);
    }
    if (st < n_full) step_full(st);
    // the last, partial step (points below nmax only): clamped reads, masked values
    if (n_full * MSTEP < nmax && n_full % MW == w) {
        const int i = n_full * MSTEP + kq * MS;
        double a[MS], bb[NT][MS];
#pragma unroll
        for (int u = 0; u < MS; ++u) {
            const int j = min(i + u, nmax - 1);
            const bool in = i + u < nmax;
            const double va = ap[j];
            a[u] = in ? va : 0.0;
#pragma unroll
            for (int c = 0; c < NT; ++c) {
                const double vb = tp[c][j];
                bb[c][u] = in ? vb : 0.0;
            }
        }
#pragma unroll
        for (int u = 0; u < MS; ++u)
#pragma unroll
            for (int c = 0; c < NT; ++c) acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u], bb[c][u], acc[c], 0, 0, 0);
    }
    if (w > 0) {
#pragma unroll
        for (int c = 0; c < NT; ++c)
#pragma unroll
            for (int v = 0; v < 4; ++v) s_red[w - 1][4 * c + v][lane] = acc[c][v];
    }
    __syncthreads();
    if (w == 0) {
#pragma unroll
        for (int q = 0; q < MW - 1; ++q)
#pragma unroll
            for (int c = 0; c < NT; ++c)
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[c][v] += s_red[q][4 * c + v][lane];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int row = kq + 4 * v;       // pair in the item
            if (row < count) {
                double* o = out + (size_t)pairs[first + row].out_index * 2 * nl;
#pragma unroll
                for (int c = 0; c < NT; ++c)
                    if (m + 16 * c < nl) o[m + 16 * c] = acc[c][v];
            }
        }
    }
}

// The same dots with the operands staged through LDS.  Loads in the MFMA
// operand layout (lane m + 16k holds row m) send 16 rows through every load
// instruction, one 16-byte access per lane: the vector L1 then spends its
// time on tag lookups (about 64 per instruction, its address path stalled
// more than half the time) rather than on bytes.  Here the workgroup and
// item are those of ddot_mfma_kernel (<= 16 pairs sharing t, four
// wavefronts taking interleaved chunks of DC points), but each wavefront
// copies its chunk's 16 alpha rows and 32 t rows straight into its own LDS
// buffers with global_load_lds_dwordx4 (a copy moves eight 128-byte rows,
// 64 lanes x 16 bytes), one chunk ahead of the one it multiplies, and reads
// the MFMA operands from there.  A copy cannot leave padding between its
// rows, so row R is stored rotated by (R >> 1) & 7 of its 16-byte columns,
// which puts the 16 rows of an operand read on distinct banks.  The last
// chunk is copied dword by dword, points at and past the item's largest n
// fetched from a zero array: nothing there is read.
constexpr int DC = 16;        // points per chunk (one 128-byte row chunk)
constexpr int DROWS = 48;     // rows per chunk buffer: 16 alpha, 32 t (multipoles 0..31)
constexpr int DCOPIES = DROWS / 2;
constexpr int DBUF = DROWS * 32;   // dwords per chunk buffer
constexpr int DNBUF = 2;           // chunk buffers per wavefront (one chunk in flight)

__device__ double g_ddot_zeros[2];

typedef __attribute__((address_space(3))) void* ddot_lds_ptr;

__device__ __forceinline__ void ddot_dma(const void* g, float* l) {
    __builtin_amdgcn_global_load_lds(g, (ddot_lds_ptr)l, 4, 0, 0);
}

template <int N>
__device__ __forceinline__ void ddot_wait() { asm volatile("s_waitcnt vmcnt(%0)" ::"n"(N) : "memory"); }

// A chunk buffer holds its rows of 128 bytes (eight 16-byte columns) with
// row R rotated by (R >> 1) & 7 columns: the 16 rows of an operand read then
// fall on distinct banks.  Double index of element e of row R:
__device__ __forceinline__ int ddot_at(int R, int e) {
    return R * 16 + ((((e >> 1) - (R >> 1)) & 7) << 1) + (e & 1);
}

template <int NT>
__global__ __launch_bounds__(64 * MW) void ddot_stage_kernel(const dwba_ddot_pair* __restrict__ pairs,
                                                             const int2* __restrict__ items, int n_items,
                                                             int L_max, double* __restrict__ out) {
    __shared__ __attribute__((aligned(16))) float s_buf[MW][DNBUF][DBUF];    // also the final reduction
    // consecutive items (one energy's pairs, sharing t) on one XCD: blocks are dealt to the 8 XCDs round-robin
    const int nb = gridDim.x, b = blockIdx.x;
    const int per = nb / 8, rem = nb % 8, xcd = b % 8;
    const int item = (xcd < rem ? xcd * (per + 1) : rem * (per + 1) + (xcd - rem) * per) + b / 8;
    if (item >= n_items) return;                           // uniform over the workgroup
    const int2 it = items[item];
    const int first = it.x, count = min(it.y, MQ);
    if (count <= 0) return;
    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int m = lane & 15, kq = lane >> 4;
    const int nl = L_max + 1;
    const double* t = pairs[first].t;
    const size_t ld = (size_t)pairs[first].ld;
    int nmax = 0;
    for (int q = 0; q < count; ++q) nmax = max(nmax, pairs[first + q].n);
    const int n_chunks = (nmax + DC - 1) / DC;
    // rows of the chunk buffer: 0..15 alpha of the item's pairs (pair 0's past the count: its results are
    // never stored), 16..47 t rows 0..31 (row 0 past L_max).  A 16-byte copy moves rows 8i .. 8i + 7 (lane
    // l: row 8i + l / 8, stored column l % 8); the last chunk, which may end inside a 16-byte column, is
    // copied dword by dword (lane l: row 2j + l / 32, stored dword l % 32), each point tested.
    const double* rows16[DROWS / 8];
    const double* rows4[DCOPIES];
    auto row_of = [&](int R) -> const double* {
        return R < 16 ? pairs[first + (R < count ? R : 0)].alpha : t + (size_t)(R - 16 < nl ? R - 16 : 0) * ld;
    };
#pragma unroll
    for (int i = 0; i < DROWS / 8; ++i) rows16[i] = row_of(8 * i + (lane >> 3));
    const void* zeros = g_ddot_zeros;
    float* mine = &s_buf[w][0][0];
    auto copy = [&](int c) {
        float* buf = mine + ((c / MW) % DNBUF) * DBUF;    // this wavefront's chunks take the buffers in turn
        if ((c + 1) * DC <= nmax) {
#pragma unroll
            for (int i = 0; i < DROWS / 8; ++i) {
                const int R = 8 * i + (lane >> 3);
                const int col = ((lane & 7) + (R >> 1)) & 7;  // the 16-byte column stored at lane & 7
                const char* src = reinterpret_cast<const char*>(rows16[i] + (size_t)c * DC) + 16 * col;
                __builtin_amdgcn_global_load_lds(src, (ddot_lds_ptr)(buf + 256 * i), 16, 0, 0);
            }
        } else {
#pragma unroll
            for (int j = 0; j < DCOPIES; ++j) {
                const int R = 2 * j + (lane >> 5);
                const int dl = lane & 31;                       // stored dword of row R
                const int e2 = (((dl >> 2) + (R >> 1)) & 7);    // its 16-byte column
                const int d = 4 * e2 + (dl & 3);                // the row's dword stored there
                const char* src = reinterpret_cast<const char*>(row_of(R) + (size_t)c * DC) + 4 * d;
                ddot_dma(c * DC + (d >> 1) < nmax ? (const void*)src : zeros, buf + 64 * j);
            }
        }
    };
    v4d acc[NT];
#pragma unroll
    for (int c = 0; c < NT; ++c) acc[c] = v4d{0.0, 0.0, 0.0, 0.0};
    if (w < n_chunks) copy(w);
    for (int c = w; c < n_chunks; c += MW) {
        if (c + MW < n_chunks) {
            copy(c + MW);
            // chunk c landed, chunk c + MW in flight (the 16-byte copies: whole chunks only, never the last)
            ddot_wait<DROWS / 8>();
        } else {
            ddot_wait<0>();
        }
        const double* buf = reinterpret_cast<const double*>(mine + ((c / MW) % DNBUF) * DBUF);
#pragma unroll
        for (int u = 0; u < DC / 4; ++u) {
            const int e = 4 * u + kq;
            const double a = buf[ddot_at(m, e)];
#pragma unroll
            for (int tc = 0; tc < NT; ++tc)
                acc[tc] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, buf[ddot_at(16 + 16 * tc + m, e)], acc[tc], 0, 0, 0);
        }
    }
    // the wavefronts' partial sums through LDS (every copy has landed and been read)
    __syncthreads();
    double* s_red = reinterpret_cast<double*>(&s_buf[0][0][0]);   // [MW - 1][4 NT][64]
    if (w > 0) {
#pragma unroll
        for (int tc = 0; tc < NT; ++tc)
#pragma unroll
            for (int v = 0; v < 4; ++v) s_red[((w - 1) * 4 * NT + 4 * tc + v) * 64 + lane] = acc[tc][v];
    }
    __syncthreads();
    if (w == 0) {
#pragma unroll
        for (int q = 0; q < MW - 1; ++q)
#pragma unroll
            for (int tc = 0; tc < NT; ++tc)
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[tc][v] += s_red[(q * 4 * NT + 4 * tc + v) * 64 + lane];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int row = kq + 4 * v;                    // pair in the item
            if (row < count) {
                double* o = out + (size_t)pairs[first + row].out_index * 2 * nl;
#pragma unroll
                for (int tc = 0; tc < NT; ++tc)
                    if (m + 16 * tc < nl) o[m + 16 * tc] = acc[tc][v];
            }
        }
    }
}

int nl_bucket_g(int L_max) {
    static const int buckets[] = {1, 2, 4, 8, 12, 16, 20, 24, 28, 32, 40, 48, 56, 64};
    for (int b : buckets)
        if (L_max + 1 <= b) return b;
    return -1;
}

}  // namespace
}  // namespace dwba

using namespace dwba;

extern "C" {

int32_t dwba_dgemv_f64(int32_t n_energies, const dwba_dgemv_energy* d_energies, int32_t max_rows,
                       int32_t L_max, void* stream) {
    DWBA_REQUIRE(initialised(), "dwba_dgemv_f64: call dwba_init first");
    DWBA_REQUIRE(n_energies > 0 && d_energies && max_rows > 0 && L_max >= 0 && L_max <= 63,
                 "dwba_dgemv_f64: bad arguments");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    dim3 grid((max_rows + GT - 1) / GT, n_energies);
    switch (nl_bucket_g(L_max)) {
#define C(n) case n: hipLaunchKernelGGL(dgemv_kernel<n>, grid, dim3(GT), 0, s, d_energies); break;
        C(1) C(2) C(4) C(8) C(12) C(16) C(20) C(24) C(28) C(32) C(40) C(48) C(56) C(64)
#undef C
        default: return set_error(DWBA_ERR_ARG, "dwba_dgemv_f64: L_max %d", L_max);
    }
    DWBA_LAUNCH_CHECK("dgemv_kernel");
    return DWBA_OK;
}

int32_t dwba_ddot_f64(int32_t n_pairs, const dwba_ddot_pair* d_pairs, int32_t L_max, double* d_out, void* stream) {
    DWBA_REQUIRE(initialised(), "dwba_ddot_f64: call dwba_init first");
    DWBA_REQUIRE(n_pairs > 0 && d_pairs && d_out && L_max >= 0 && L_max <= 63, "dwba_ddot_f64: bad arguments");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    switch (nl_bucket_g(L_max)) {
#define C(n) case n: hipLaunchKernelGGL(ddot_kernel<n>, dim3(n_pairs), dim3(256), 0, s, d_pairs, L_max, d_out); break;
        C(1) C(2) C(4) C(8) C(12) C(16) C(20) C(24) C(28) C(32) C(40) C(48) C(56) C(64)
#undef C
        default: return set_error(DWBA_ERR_ARG, "dwba_ddot_f64: L_max %d", L_max);
    }
    DWBA_LAUNCH_CHECK("ddot_kernel");
    return DWBA_OK;
}

int32_t dwba_ddot_tiled_f64(int32_t n_items, const int32_t* d_items, const dwba_ddot_pair* d_pairs, int32_t L_max,
                            double* d_out, void* stream) {
    DWBA_REQUIRE(initialised(), "dwba_ddot_tiled_f64: call dwba_init first");
    DWBA_REQUIRE(n_items > 0 && d_items && d_pairs && d_out && L_max >= 0 && L_max <= 31,
                 "dwba_ddot_tiled_f64: bad arguments (L_max <= 31)");
    hipLaunchKernelGGL(ddot_tiled_kernel, dim3(n_items), dim3(256), 0, reinterpret_cast<hipStream_t>(stream), d_pairs,
                       reinterpret_cast<const int2*>(d_items), L_max, d_out);
    DWBA_LAUNCH_CHECK("ddot_tiled_kernel");
    return DWBA_OK;
}

int32_t dwba_ddot_mfma_f64(int32_t n_items, const int32_t* d_items, const dwba_ddot_pair* d_pairs, int32_t L_max,
                           double* d_out, void* stream) {
    DWBA_REQUIRE(initialised(), "dwba_ddot_mfma_f64: call dwba_init first");
    DWBA_REQUIRE(n_items > 0 && d_items && d_pairs && d_out && L_max >= 0 && L_max <= 31,
                 "dwba_ddot_mfma_f64: bad arguments (L_max <= 31)");
    if (L_max < 16)
        hipLaunchKernelGGL(ddot_mfma_kernel<1>, dim3(n_items), dim3(64 * MW), 0,
                           reinterpret_cast<hipStream_t>(stream), d_pairs, reinterpret_cast<const int2*>(d_items),
                           n_items, L_max, d_out);
    else
        hipLaunchKernelGGL(ddot_mfma_kernel<2>, dim3(n_items), dim3(64 * MW), 0,
                           reinterpret_cast<hipStream_t>(stream), d_pairs, reinterpret_cast<const int2*>(d_items),
                           n_items, L_max, d_out);
    DWBA_LAUNCH_CHECK("ddot_mfma_kernel");
    return DWBA_OK;
}

int32_t dwba_ddot_stage_f64(int32_t n_items, const int32_t* d_items, const dwba_ddot_pair* d_pairs, int32_t L_max,
                            double* d_out, void* stream) {
    DWBA_REQUIRE(initialised(), "dwba_ddot_stage_f64: call dwba_init first");
    DWBA_REQUIRE(n_items > 0 && d_items && d_pairs && d_out && L_max >= 0 && L_max <= 31,
                 "dwba_ddot_stage_f64: bad arguments (L_max <= 31)");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (L_max < 16)
        hipLaunchKernelGGL(ddot_stage_kernel<1>, dim3(n_items), dim3(64 * MW), 0, s, d_pairs,
                           reinterpret_cast<const int2*>(d_items), n_items, L_max, d_out);
    else
        hipLaunchKernelGGL(ddot_stage_kernel<2>, dim3(n_items), dim3(64 * MW), 0, s, d_pairs,
                           reinterpret_cast<const int2*>(d_items), n_items, L_max, d_out);
    DWBA_LAUNCH_CHECK("ddot_stage_kernel");
    return DWBA_OK;
}

int32_t dwba_batch_finalize(int32_t n_pairs, const dwba_finalize_pair* d_pairs, int32_t L_max, double* d_out,
                            void* stream) {
    DWBA_REQUIRE(initialised(), "dwba_batch_finalize: call dwba_init first");
    DWBA_REQUIRE(n_pairs > 0 && d_pairs && d_out && L_max >= 0, "dwba_batch_finalize: bad arguments");
    hipLaunchKernelGGL(finalize_kernel, dim3((n_pairs + 255) / 256), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), d_pairs, n_pairs, L_max, d_out);
    DWBA_LAUNCH_CHECK("finalize_kernel");
    return DWBA_OK;
}

}  // extern "C"
