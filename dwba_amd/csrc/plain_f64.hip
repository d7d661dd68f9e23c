// il_plain_f64: I_L = alpha^T K_L beta for all L, direct + exchange, FP64.
//
// Reference work replaced: the per-L dense kernel build and the two GEMVs of
// radial_ME_all_L (dwba_matrix_elements.py:1057-1080, :1113-1116,
// :1243-1245, :1288-1289) and of its CuPy twin
// (dwba_matrix_elements.py:2318-2515).  K_L is never stored:
//
//  * K is symmetric, so the kernel walks the upper triangle of 64x64 tiles
//    and folds the lower half into the coefficient
//        c_ij = alpha_i beta_j + alpha_j beta_i      (i < j),  c_ii = alpha_i beta_i;
//  * for i < j, K_0 = 1/r_j and K_{L+1} = K_L * min(r_i / r_j, 1 - 1e-12); the
//    diagonal uses K_0 = 1/r_i and ratio (1 - 1e-12), matching the reference's
//    clamp inv_gtr * exp(L log(min(ratio, 1-1e-12) + 1e-300));
//  * per element and L: one multiply and two FMAs (direct, exchange); each
//    thread keeps 2*NL FP64 accumulators across all the tiles of its block.
//
// Work mapping: a 256-thread block owns a contiguous run of `tpb` tiles of one
// job; in a tile, lane = row (64 rows) and wave w owns columns 16w..16w+15, so
// column data are LDS broadcasts.  Blocks write 2*NL partial sums; a second
// kernel adds them per job in a fixed order (bit-reproducible results).
#include "common.hpp"

#include <atomic>

// Device scratch of dwba_il_plain_f64_noalloc (size in include/dwba_hip.h);
// each GPU holds its own copy until the process ends.
static __device__ __align__(256) unsigned char scratch[DWBA_NOALLOC_SCRATCH_BYTES];

namespace dwba {
namespace {

constexpr int PT = 64;        // tile edge
constexpr int PTHREADS = 256; // 4 wavefronts
constexpr double KCLAMP = 1.0 - 1e-12;

__host__ __device__ inline int64_t tri_tiles(int64_t tr, int64_t tc) {
    // upper-triangle tiles (ti < tr, ti <= tj < tc), tr <= tc
    return tr * tc - tr * (tr - 1) / 2;
}

__host__ __device__ inline int tiles_of(int n) { return (n + PT - 1) / PT; }

template <int NL>
__global__ __launch_bounds__(PTHREADS, (NL > 48 ? 1 : (NL > 16 ? 2 : 3))) void plain_f64_kernel(const dwba_plain_job* __restrict__ jobs,
                                                             const int* __restrict__ block_off,
                                                             int n_jobs, int tpb,
                                                             double* __restrict__ partial) {
    __shared__ double s_row[5][PT];  // r, ad, ae, bd, be
    __shared__ double s_col[6][PT];  // r, 1/r, ad, ae, bd, be
    __shared__ double s_red[4][2 * NL];

    const int b = blockIdx.x;
    int lo = 0, hi = n_jobs - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (block_off[mid] <= b) lo = mid; else hi = mid - 1;
    }
    const dwba_plain_job J = jobs[lo];
    const int lb = b - block_off[lo];
    const int tr = tiles_of(J.n_rows), tc = tiles_of(J.n_cols);
    const int64_t ntiles = tri_tiles(tr, tc);
    int64_t t = (int64_t)lb * tpb;
    const int64_t t_end = t + tpb < ntiles ? t + tpb : ntiles;

    // (ti, tj) of the first tile: rows before ti hold tri_tiles(ti, tc) tiles.
    int ti = 0;
    while (ti + 1 < tr && tri_tiles(ti + 1, tc) <= t) ++ti;
    int tj = ti + (int)(t - tri_tiles(ti, tc));

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;

    double acc_d[NL], acc_e[NL];
#pragma unroll
    for (int L = 0; L < NL; ++L) { acc_d[L] = 0.0; acc_e[L] = 0.0; }

    for (; t < t_end; ++t) {
        const int i0 = ti * PT, j0 = tj * PT;
        __syncthreads();
        if (threadIdx.x < PT) {
            const int i = i0 + threadIdx.x;
            const bool in_c = i < J.n_cols, in_r = i < J.n_rows;
            s_row[0][threadIdx.x] = in_c ? J.r[i] : 1.0;
            s_row[1][threadIdx.x] = in_r ? J.alpha_dir[i] : 0.0;
            s_row[2][threadIdx.x] = in_r ? J.alpha_ex[i] : 0.0;
            s_row[3][threadIdx.x] = in_c ? J.beta_dir[i] : 0.0;
            s_row[4][threadIdx.x] = in_c ? J.beta_ex[i] : 0.0;
        } else if (threadIdx.x < 2 * PT) {
            const int c = threadIdx.x - PT;
            const int j = j0 + c;
            const bool in_c = j < J.n_cols, in_r = j < J.n_rows;
            const double rj = in_c ? J.r[j] : 1.0;
            s_col[0][c] = rj;
            s_col[1][c] = 1.0 / rj;
            s_col[2][c] = in_r ? J.alpha_dir[j] : 0.0;
            s_col[3][c] = in_r ? J.alpha_ex[j] : 0.0;
            s_col[4][c] = in_c ? J.beta_dir[j] : 0.0;
            s_col[5][c] = in_c ? J.beta_ex[j] : 0.0;
        }
        __syncthreads();
        const double ri = s_row[0][lane];
        const double adi = s_row[1][lane], aei = s_row[2][lane];
        const double bdi = s_row[3][lane], bei = s_row[4][lane];
        const bool dtile = (ti == tj);
        const double iri = s_col[1][lane];
        // two columns per step: two independent K chains per lane
#pragma unroll 1
        for (int k = 0; k < 16; k += 2) {
            const int c0 = wave * 16 + k, c1 = c0 + 1;
            const double ir0 = s_col[1][c0], ir1 = s_col[1][c1];
            double cd0 = fma(adi, s_col[4][c0], s_col[2][c0] * bdi);
            double ce0 = fma(aei, s_col[5][c0], s_col[3][c0] * bei);
            double cd1 = fma(adi, s_col[4][c1], s_col[2][c1] * bdi);
            double ce1 = fma(aei, s_col[5][c1], s_col[3][c1] * bei);
            double q0 = fmin(ri * ir0, KCLAMP), q1 = fmin(ri * ir1, KCLAMP);
            double K0 = ir0, K1 = ir1;
            if (dtile) {  // in a diagonal tile: drop i > j, special-case i == j
                if (lane > c0) { cd0 = 0.0; ce0 = 0.0; }
                else if (lane == c0) { cd0 = adi * bdi; ce0 = aei * bei; q0 = KCLAMP; K0 = iri; }
                if (lane > c1) { cd1 = 0.0; ce1 = 0.0; }
                else if (lane == c1) { cd1 = adi * bdi; ce1 = aei * bei; q1 = KCLAMP; K1 = iri; }
            }
#pragma unroll
            for (int L = 0; L < NL; ++L) {
                acc_d[L] = fma(cd0, K0, acc_d[L]);
                acc_e[L] = fma(ce0, K0, acc_e[L]);
                acc_d[L] = fma(cd1, K1, acc_d[L]);
                acc_e[L] = fma(ce1, K1, acc_e[L]);
                // keep the power chain in step with the accumulation so the
                // scheduler does not materialise all K_L at once
                asm volatile("" : "+v"(K0), "+v"(K1), "+v"(acc_d[L]), "+v"(acc_e[L]));
                K0 *= q0;
                K1 *= q1;
            }
        }
        if (++tj == tc) { ++ti; tj = ti; }
    }

    // block reduction: butterfly within each wave, then the 4 waves in order
#pragma unroll
    for (int L = 0; L < NL; ++L) {
        const double d = wave_sum(acc_d[L]);
        const double e = wave_sum(acc_e[L]);
        if (lane == 0) { s_red[wave][L] = d; s_red[wave][NL + L] = e; }
        __builtin_amdgcn_sched_barrier(0);  // one L at a time: bounded register pressure
    }
    __syncthreads();
    for (int x = threadIdx.x; x < 2 * NL; x += PTHREADS)
        partial[(size_t)b * 2 * NL + x] = ((s_red[0][x] + s_red[1][x]) + s_red[2][x]) + s_red[3][x];
}

// Sum the partials of each job: one 256-thread block per (job, output value);
// thread t adds blocks b0+t, b0+t+256, ... and the block reduces in a fixed
// tree order (bit-reproducible).  out is [job][2][L_max+1].
__global__ __launch_bounds__(256) void plain_f64_finalize(const int* __restrict__ block_off, int NL, int L_max,
                                                          const double* __restrict__ partial,
                                                          double* __restrict__ out) {
    __shared__ double s_red[4];
    const int k = blockIdx.x;
    const int nl = L_max + 1;
    const int x = blockIdx.y;                 // 0 .. 2*nl-1
    const int fam = x / nl, L = x - fam * nl;
    const int src = fam * NL + L;
    const int b0 = block_off[k], b1 = block_off[k + 1];
    double s = 0.0;
    for (int b = b0 + (int)threadIdx.x; b < b1; b += 256) s += partial[(size_t)b * 2 * NL + src];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[(size_t)k * 2 * nl + x] = ((s_red[0] + s_red[1]) + s_red[2]) + s_red[3];
}

// Per-job block counts -> exclusive prefix sums block_off[0..n_jobs].
__global__ void plain_f64_plan(const dwba_plain_job* __restrict__ jobs, int n_jobs, int tpb,
                               int* __restrict__ block_off) {
    __shared__ int s_tot[1024];
    const int nt = blockDim.x;
    const int per = (n_jobs + nt - 1) / nt;
    const int k0 = threadIdx.x * per;
    const int k1 = k0 + per < n_jobs ? k0 + per : n_jobs;
    int sum = 0;
    for (int k = k0; k < k1; ++k) {
        const int64_t nt_k = tri_tiles(tiles_of(jobs[k].n_rows), tiles_of(jobs[k].n_cols));
        sum += (int)((nt_k + tpb - 1) / tpb);
    }
    s_tot[threadIdx.x] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        int run = 0;
        for (int x = 0; x < nt; ++x) { const int v = s_tot[x]; s_tot[x] = run; run += v; }
        block_off[n_jobs] = run;
    }
    __syncthreads();
    int run = s_tot[threadIdx.x];
    for (int k = k0; k < k1; ++k) {
        block_off[k] = run;
        const int64_t nt_k = tri_tiles(tiles_of(jobs[k].n_rows), tiles_of(jobs[k].n_cols));
        run += (int)((nt_k + tpb - 1) / tpb);
    }
}

__global__ void plain_f64_store_job(dwba_plain_job job, dwba_plain_job* dst) {
    if (threadIdx.x == 0) *dst = job;
}

int nl_bucket(int L_max) {
    static const int buckets[] = {1, 2, 4, 8, 12, 16, 20, 24, 28, 32, 40, 48, 56, 64};
    for (int b : buckets)
        if (L_max + 1 <= b) return b;
    return -1;
}

struct PlainPlan {
    int NL = 0;
    int tpb = 1;
    int n_blocks = 0;
    size_t off_bytes = 0, partial_bytes = 0, job_bytes = 0;
    size_t total() const { return off_bytes + partial_bytes + job_bytes; }
};

PlainPlan make_plan(int n_jobs, const int32_t* n_rows, const int32_t* n_cols, int L_max) {
    PlainPlan p;
    p.NL = nl_bucket(L_max);
    int64_t tot = 0;
    for (int k = 0; k < n_jobs; ++k) tot += tri_tiles(tiles_of(n_rows[k]), tiles_of(n_cols[k]));
    // ~4 resident blocks per CU, 4 rounds of them, but never fewer than 1 tile per block
    const int64_t target = (int64_t)device_cu_count() * 16;
    p.tpb = (int)std::max<int64_t>(1, (tot + target - 1) / target);
    int64_t nb = 0;
    for (int k = 0; k < n_jobs; ++k) {
        const int64_t t = tri_tiles(tiles_of(n_rows[k]), tiles_of(n_cols[k]));
        nb += (t + p.tpb - 1) / p.tpb;
    }
    p.n_blocks = (int)nb;
    p.off_bytes = align_up(sizeof(int) * (size_t)(n_jobs + 1));
    p.partial_bytes = align_up(sizeof(double) * 2 * (size_t)(p.NL > 0 ? p.NL : 1) * (size_t)nb);
    p.job_bytes = align_up(sizeof(dwba_plain_job));
    return p;
}

template <int NL>
void launch_nl(int nb, hipStream_t s, const dwba_plain_job* jobs, const int* off, int n_jobs, int tpb,
               double* partial) {
    hipLaunchKernelGGL(plain_f64_kernel<NL>, dim3(nb), dim3(PTHREADS), 0, s, jobs, off, n_jobs, tpb,
                       partial);
}

int launch_plain(const PlainPlan& p, int n_jobs, const dwba_plain_job* d_jobs, int L_max, double* d_out,
                 char* ws, hipStream_t s) {
    int* d_off = reinterpret_cast<int*>(ws);
    double* d_part = reinterpret_cast<double*>(ws + p.off_bytes);
    hipLaunchKernelGGL(plain_f64_plan, dim3(1), dim3(1024), 0, s, d_jobs, n_jobs, p.tpb, d_off);
    DWBA_LAUNCH_CHECK("plain_f64_plan");
    if (p.n_blocks > 0) {
        switch (p.NL) {
#define DWBA_NL_CASE(n) case n: launch_nl<n>(p.n_blocks, s, d_jobs, d_off, n_jobs, p.tpb, d_part); break;
            DWBA_NL_CASE(1) DWBA_NL_CASE(2) DWBA_NL_CASE(4) DWBA_NL_CASE(8) DWBA_NL_CASE(12)
            DWBA_NL_CASE(16) DWBA_NL_CASE(20) DWBA_NL_CASE(24) DWBA_NL_CASE(28) DWBA_NL_CASE(32)
            DWBA_NL_CASE(40) DWBA_NL_CASE(48) DWBA_NL_CASE(56) DWBA_NL_CASE(64)
#undef DWBA_NL_CASE
            default: return set_error(DWBA_ERR_ARG, "il_plain_f64: unsupported L_max %d", L_max);
        }
        DWBA_LAUNCH_CHECK("plain_f64_kernel");
    }
    hipLaunchKernelGGL(plain_f64_finalize, dim3(n_jobs, 2 * (L_max + 1)), dim3(256), 0, s, d_off, p.NL, L_max,
                       d_part, d_out);
    DWBA_LAUNCH_CHECK("plain_f64_finalize");
    return DWBA_OK;
}

// One job through the plain kernels with the caller's workspace `ws` of
// `ws_bytes` bytes: the job descriptor is stored into the workspace by a
// kernel (no host-device copy), so the call can be captured in a HIP graph.
int plain_single(const char* name, const double* d_r, int32_t n_rows, int32_t n_cols, const double* d_alpha_dir,
                 const double* d_beta_dir, const double* d_alpha_ex, const double* d_beta_ex, int32_t L_max,
                 double* d_out, char* ws, size_t ws_bytes, hipStream_t s) {
    DWBA_REQUIRE(initialised(), "%s: call dwba_init first", name);
    DWBA_REQUIRE(L_max >= 0 && L_max <= 63, "%s: L_max=%d outside [0, 63]", name, L_max);
    DWBA_REQUIRE(n_rows >= 0 && n_cols >= n_rows, "%s: need 0 <= n_rows <= n_cols (got %d, %d)", name, n_rows,
                 n_cols);
    DWBA_REQUIRE(d_out != nullptr, "%s: null output", name);
    DWBA_REQUIRE(n_cols == 0 || (d_r && d_alpha_dir && d_beta_dir && d_alpha_ex && d_beta_ex),
                 "%s: null input pointer", name);
    PlainPlan p = make_plan(1, &n_rows, &n_cols, L_max);
    if (ws_bytes < p.total())
        return set_error(DWBA_ERR_WORKSPACE, "%s: workspace %zu < %zu bytes", name, ws_bytes, p.total());
    dwba_plain_job* d_job = reinterpret_cast<dwba_plain_job*>(ws + p.off_bytes + p.partial_bytes);
    dwba_plain_job job{d_r, d_alpha_dir, d_beta_dir, d_alpha_ex, d_beta_ex, n_rows, n_cols};
    hipLaunchKernelGGL(plain_f64_store_job, dim3(1), dim3(64), 0, s, job, d_job);
    DWBA_LAUNCH_CHECK("plain_f64_store_job");
    return launch_plain(p, 1, d_job, L_max, d_out, ws, s);
}

// Scratch of dwba_il_plain_f64_noalloc.  `scratch` names the array in device
// code only; on the host it is a stand-in symbol, so host code works with the
// device address that hipGetSymbolAddress returns, recorded per device by
// noalloc_scratch_setup (called from dwba_init with that device current).
// For one job the plain kernels use at most 256 B of block offsets, 2*64
// doubles per workgroup for <= 16 workgroups per CU, and one job descriptor:
// about 4 MiB on 256 CUs.  Nothing in it needs to start at zero (the plan
// kernel writes every offset, each workgroup every one of its partial sums).
constexpr int kMaxScratchDevices = 64;
std::atomic<void*> g_scratch_ptr[kMaxScratchDevices];

}  // namespace

int noalloc_scratch_setup(int device) {
    if (device < 0 || device >= kMaxScratchDevices)
        return set_error(DWBA_ERR_ARG, "dwba_init: device %d beyond the %d devices the scratch table covers", device,
                         kMaxScratchDevices);
    if (g_scratch_ptr[device].load() != nullptr) return DWBA_OK;
    void* p = nullptr;
    DWBA_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (p == nullptr) return set_error(DWBA_ERR_HIP, "dwba_init: hipGetSymbolAddress(scratch) returned NULL");
    g_scratch_ptr[device].store(p);
    return DWBA_OK;
}

}  // namespace dwba

using namespace dwba;

extern "C" {

size_t dwba_il_plain_f64_workspace_bytes(int32_t n_rows, int32_t n_cols, int32_t L_max) {
    if (n_rows < 0 || n_cols < n_rows || L_max < 0 || L_max > 63) return 0;
    return make_plan(1, &n_rows, &n_cols, L_max).total();
}

int32_t dwba_il_plain_f64(const double* d_r, int32_t n_rows, int32_t n_cols, const double* d_alpha_dir,
                          const double* d_beta_dir, const double* d_alpha_ex, const double* d_beta_ex,
                          int32_t L_max, double* d_out, void* d_workspace, size_t workspace_bytes,
                          void* stream) {
    DWBA_REQUIRE(d_workspace != nullptr, "dwba_il_plain_f64: null workspace");
    return plain_single("dwba_il_plain_f64", d_r, n_rows, n_cols, d_alpha_dir, d_beta_dir, d_alpha_ex, d_beta_ex,
                        L_max, d_out, reinterpret_cast<char*>(d_workspace), workspace_bytes,
                        reinterpret_cast<hipStream_t>(stream));
}

int32_t dwba_il_plain_f64_noalloc(const double* d_r, int32_t n_rows, int32_t n_cols, const double* d_alpha_dir,
                                  const double* d_beta_dir, const double* d_alpha_ex, const double* d_beta_ex,
                                  int32_t L_max, double* d_out, void* stream) {
    int dev = -1;
    DWBA_HIP_CHECK(hipGetDevice(&dev));
    char* base = (dev >= 0 && dev < kMaxScratchDevices) ? static_cast<char*>(g_scratch_ptr[dev].load()) : nullptr;
    DWBA_REQUIRE(base != nullptr, "dwba_il_plain_f64_noalloc: no scratch address for device %d (call dwba_init(%d))",
                 dev, dev);
    return plain_single("dwba_il_plain_f64_noalloc", d_r, n_rows, n_cols, d_alpha_dir, d_beta_dir, d_alpha_ex,
                        d_beta_ex, L_max, d_out, base, DWBA_NOALLOC_SCRATCH_BYTES,
                        reinterpret_cast<hipStream_t>(stream));
}

size_t dwba_il_plain_f64_batch_workspace_bytes(int32_t n_jobs, const int32_t* h_n_rows,
                                               const int32_t* h_n_cols, int32_t L_max) {
    if (n_jobs <= 0 || !h_n_rows || !h_n_cols || L_max < 0 || L_max > 63) return 0;
    return make_plan(n_jobs, h_n_rows, h_n_cols, L_max).total();
}

int32_t dwba_il_plain_f64_batch(int32_t n_jobs, const dwba_plain_job* d_jobs, const int32_t* h_n_rows,
                                const int32_t* h_n_cols, int32_t L_max, double* d_out, void* d_workspace,
                                size_t workspace_bytes, void* stream) {
    DWBA_REQUIRE(initialised(), "dwba_il_plain_f64_batch: call dwba_init first");
    DWBA_REQUIRE(n_jobs > 0 && d_jobs && h_n_rows && h_n_cols && d_out && d_workspace,
                 "dwba_il_plain_f64_batch: bad arguments");
    DWBA_REQUIRE(L_max >= 0 && L_max <= 63, "dwba_il_plain_f64_batch: L_max=%d outside [0, 63]", L_max);
    for (int k = 0; k < n_jobs; ++k)
        DWBA_REQUIRE(h_n_rows[k] >= 0 && h_n_cols[k] >= h_n_rows[k],
                     "dwba_il_plain_f64_batch: job %d needs 0 <= n_rows <= n_cols", k);
    PlainPlan p = make_plan(n_jobs, h_n_rows, h_n_cols, L_max);
    if (workspace_bytes < p.total())
        return set_error(DWBA_ERR_WORKSPACE, "dwba_il_plain_f64_batch: workspace %zu < %zu bytes",
                         workspace_bytes, p.total());
    return launch_plain(p, n_jobs, d_jobs, L_max, d_out, reinterpret_cast<char*>(d_workspace),
                        reinterpret_cast<hipStream_t>(stream));
}

}  // extern "C"
