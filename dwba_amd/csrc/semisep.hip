// il_semisep_f64: I_L = alpha^T K_L beta for all L, direct + exchange, FP64,
// in O(N) work per (pair, L) instead of O(N^2).
//
// Same jobs, output layout and meaning as dwba_il_plain_f64_batch (the
// per-L kernel build + GEMVs of radial_ME_all_L, dwba_matrix_elements.py:
// 1057-1080, :1113-1116, :1243-1245, :1281-1289); only the order of the FP64
// arithmetic differs.  K_L is semiseparable on an increasing grid:
//     K_L(i, j) = (r_j / r_i)^L / r_i  (j < i),   (r_i / r_j)^L / r_j  (j > i),
//     K_L(i, i) = (1 - 1e-12)^L / r_i          (the reference's clamp),
// so with the forward recurrences  (q_i = r_i / r_{i+1} < 1)
//     A_{i+1} = q_i^L (A_i + beta_i),      A_i  = sum_{j<i} (r_j / r_i)^L beta_j
//     Z_{i+1} = q_i^L (Z_i + alpha_i),     Z_i  = sum_{j<i} (r_j / r_i)^L alpha_j
// the whole double sum is one pass over the grid:
//     I_L = sum_i (1 / r_i) (alpha_i A_i + beta_i Z_i + d_L alpha_i beta_i).
// Every multiplier is below one, so rounding errors are damped, not amplified.
//
// Parallel form: a task is one wavefront over a contiguous range of the grid
// of one (job, family); each lane runs the recurrences from zero over its own
// contiguous sub-range and also accumulates
//     Ca = sum alpha_i / r_i * P_i,  Cb = sum beta_i / r_i * P_i,  P_i = (r_a / r_i)^L
// (r_a: first point of the lane), which is what a carried-in A or Z adds to the
// lane's sum.  The lanes are joined by an in-wave scan of the affine maps
// (M, A) -> (M1 M2, A1 M2 + A2), M = (r_a / r_b)^L; tasks of one job are joined
// in order by a second kernel.  L is processed in blocks of LB values held in
// registers (6 * LB doubles per lane); each block re-reads the lane's range.
//
// Other forms of the same contraction in this file:
//   semisep_lane_kernel        one lane per (job, family, L-block) over the whole range;
//   semisep_lane_tab_kernel    the same reading 1/r, q, q^(LB b) from per-grid tables
//                              (semisep_lane_table_kernel), jobs in any order via an output index;
//   semisep_lane_split_kernel  each range split between two wavefronts joined through LDS;
//   semisep_gemv_scan_kernel   t_L = K_L beta for one shared beta per energy (chunked affine
//                              scans), for the direct family of pairs that share it.
#include "common.hpp"

namespace dwba {
namespace {

constexpr int ST = 256;          // threads per workgroup (4 tasks)
constexpr int NSUM = 6;          // per (task, L): total, Ca, Cb, A_out, Z_out, M
constexpr int GU = 8;            // grid points per load group

__device__ inline int find_job(const int* __restrict__ off, int n_jobs, int t) {
    int lo = 0, hi = n_jobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= t) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// x^e, e >= 1, by squaring (runtime exponent).
__device__ inline double pow_int_rt(double x, int e) {
    double r = 1.0;
    while (e > 0) {
        if (e & 1) r *= x;
        x *= x;
        e >>= 1;
    }
    return r;
}

// q^L0 for L0 a multiple of LB (LB a power of two): LB-th power by squaring,
// then L0/LB - 1 multiplications; every trip count is uniform (no selects).
template <int LB>
__device__ inline double pow_block(double q, int L0) {
    if (L0 == 0) return 1.0;
    double qb = q;
#pragma unroll
    for (int m = 1; m < LB; m <<= 1) qb *= qb;
    double r = qb;
    for (int k = LB; k < L0; k += LB) r *= qb;
    return r;
}

// Tasks of job k: 2 families x W_k range tasks, W_k = ceil(n_cols / (64 * seg)).
__device__ inline int tasks_of(int n_cols, int seg) {
    if (n_cols <= 0) return 0;
    const int span = 64 * seg;
    return 2 * ((n_cols + span - 1) / span);
}

// One grid point of the recurrences for LB multipoles L0..L0+LB-1.
// HAS_A: alpha_i may be non-zero (i < n_rows); beyond the rows only the beta
// recurrence and the decay of Z remain.
template <int LB, bool HAS_A>
__device__ __forceinline__ void ss_point(double ri, double r_next, double a_i, double b_i, int L0, double& diag,
                                         double* A, double* Z, double* acc, double* Ca, double* Cb, double* P,
                                         double& ir_cur) {
    const double iri = ir_cur;
    const double ir_next = 1.0 / r_next;
    const double q = ri * ir_next;
    const double bw = b_i * iri;
    double qL = pow_block<LB>(q, L0);
    if (HAS_A) {
        const double aw = a_i * iri;
        diag = fma(a_i, bw, diag);             // the diagonal term factorises: d_L sum_i alpha_i beta_i / r_i
#pragma unroll
        for (int l = 0; l < LB; ++l) {
            acc[l] = fma(aw, A[l], acc[l]);
            acc[l] = fma(bw, Z[l], acc[l]);
            Ca[l] = fma(aw, P[l], Ca[l]);
            Cb[l] = fma(bw, P[l], Cb[l]);
            A[l] = qL * (A[l] + b_i);
            Z[l] = qL * (Z[l] + a_i);
            P[l] *= qL;
            qL *= q;
        }
    } else {
#pragma unroll
        for (int l = 0; l < LB; ++l) {
            acc[l] = fma(bw, Z[l], acc[l]);
            Cb[l] = fma(bw, P[l], Cb[l]);
            A[l] = qL * (A[l] + b_i);
            Z[l] *= qL;
            P[l] *= qL;
            qL *= q;
        }
    }
    ir_cur = ir_next;
}

// Points [i0, i1) of one lane: groups of GU with all loads issued first,
// then single points.
template <int LB, bool HAS_A>
__device__ __forceinline__ void ss_range(const double* __restrict__ r, const double* __restrict__ al,
                                         const double* __restrict__ be, int n, int i0, int i1, int L0,
                                         double& diag, double* A, double* Z, double* acc, double* Ca, double* Cb,
                                         double* P, double& ir_cur) {
    int i = i0;
#pragma unroll 1
    for (; i + GU <= i1; i += GU) {
        double rr[GU + 1], av[GU], bv[GU];
        rr[0] = r[i];
#pragma unroll
        for (int u = 0; u < GU; ++u) {
            rr[u + 1] = r[min(i + u + 1, n - 1)];
            av[u] = HAS_A ? al[i + u] : 0.0;
            bv[u] = be[i + u];
        }
#pragma unroll
        for (int u = 0; u < GU; ++u)
            ss_point<LB, HAS_A>(rr[u], rr[u + 1], av[u], bv[u], L0, diag, A, Z, acc, Ca, Cb, P, ir_cur);
    }
#pragma unroll 1
    for (; i < i1; ++i)
        ss_point<LB, HAS_A>(r[i], r[min(i + 1, n - 1)], HAS_A ? al[i] : 0.0, be[i], L0, diag, A, Z, acc, Ca, Cb, P,
                            ir_cur);
}

template <int LB, int NB>
__global__ __launch_bounds__(ST) void semisep_kernel(const dwba_plain_job* __restrict__ jobs,
                                                     const int* __restrict__ task_off, int n_jobs, int seg,
                                                     int L_max, double* __restrict__ summ) {
    const int lane = threadIdx.x & 63;
    const int t = blockIdx.x * (ST / 64) + (threadIdx.x >> 6);
    if (t >= task_off[n_jobs]) return;
    const int k = find_job(task_off, n_jobs, t);
    const dwba_plain_job J = jobs[k];
    const int local = t - task_off[k];
    const int w_per_fam = tasks_of(J.n_cols, seg) / 2;
    const int fam = local / w_per_fam;
    const int wt = local - fam * w_per_fam;
    if (J.n_rows <= 0) return;            // the join writes 0 for such jobs
    if (fam == 0 && J.alpha_dir == nullptr) return;   // direct family computed elsewhere
    const double* __restrict__ al = fam ? J.alpha_ex : J.alpha_dir;
    const double* __restrict__ be = fam ? J.beta_ex : J.beta_dir;
    const int n = J.n_cols;
    const int nr = J.n_rows;
    const int span = 64 * seg;
    const int e0 = wt * span;
    const int e1 = min(n, e0 + span);
    const int sl = (e1 - e0 + 63) / 64;
    const int a = min(e1, e0 + lane * sl);
    const int b = min(e1, a + sl);
    const int NL = L_max + 1;
    const double ldl = log(1.0 - 1e-12);
    double* __restrict__ out = summ + (size_t)t * NL * NSUM;

#pragma unroll
    for (int blk = 0; blk < NB; ++blk) {
        const int L0 = blk * LB;             // a constant after unrolling: q^L0 is straight-line code
        double A[LB], Z[LB], acc[LB], Ca[LB], Cb[LB], P[LB];
#pragma unroll
        for (int l = 0; l < LB; ++l) {
            A[l] = 0.0; Z[l] = 0.0; acc[l] = 0.0; Ca[l] = 0.0; Cb[l] = 0.0; P[l] = 1.0;
        }
        double diag = 0.0;
        double ir_cur = 1.0 / J.r[min(a, n - 1)];
        const int am = min(b, max(a, nr));               // [a, am): alpha may be non-zero
        ss_range<LB, true>(J.r, al, be, n, a, am, L0, diag, A, Z, acc, Ca, Cb, P, ir_cur);
        ss_range<LB, false>(J.r, al, be, n, am, b, L0, diag, A, Z, acc, Ca, Cb, P, ir_cur);
        const double dsum = wave_sum(diag);   // this task's sum_i alpha_i beta_i / r_i
        // join the lanes: inclusive scan of (M, A, Z) over lanes, M = P at lane end
#pragma unroll
        for (int l = 0; l < LB; ++l) {
            if (L0 + l >= NL) break;
            double M = P[l], Ai = A[l], Zi = Z[l];
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const double Mp = __shfl_up(M, d, 64);
                const double Ap = __shfl_up(Ai, d, 64);
                const double Zp = __shfl_up(Zi, d, 64);
                if (lane >= d) {
                    Ai = fma(Ap, M, Ai);
                    Zi = fma(Zp, M, Zi);
                    M = Mp * M;
                }
            }
            // exclusive prefix: carry into this lane and the factor from the task start
            double Ain = __shfl_up(Ai, 1, 64), Zin = __shfl_up(Zi, 1, 64), Mpre = __shfl_up(M, 1, 64);
            if (lane == 0) { Ain = 0.0; Zin = 0.0; Mpre = 1.0; }
            const int L = L0 + l;
            const double tot = wave_sum(acc[l] + Ain * Ca[l] + Zin * Cb[l]) +
                               (L == 0 ? 1.0 : exp((double)L * ldl)) * dsum;
            const double ca = wave_sum(Mpre * Ca[l]);
            const double cb = wave_sum(Mpre * Cb[l]);
            const double A_out = __shfl(Ai, 63, 64), Z_out = __shfl(Zi, 63, 64), M_out = __shfl(M, 63, 64);
            if (lane == 0) {
                double* o = out + (size_t)(L0 + l) * NSUM;
                o[0] = tot; o[1] = ca; o[2] = cb; o[3] = A_out; o[4] = Z_out; o[5] = M_out;
            }
        }
    }
}

// Join the tasks of each (job, family) in grid order: one thread per L.
__global__ void semisep_join(const dwba_plain_job* __restrict__ jobs, const int* __restrict__ task_off, int seg,
                             int L_max, const double* __restrict__ summ, double* __restrict__ out) {
    const int k = blockIdx.x, fam = blockIdx.y, L = threadIdx.x;
    const int NL = L_max + 1;
    if (L >= NL) return;
    if (fam == 0 && jobs[k].alpha_dir == nullptr) return;   // left to the caller
    const int w = tasks_of(jobs[k].n_cols, seg) / 2;
    double tot = 0.0, Ain = 0.0, Zin = 0.0;
    if (jobs[k].n_rows > 0) {
        for (int x = 0; x < w; ++x) {
            const double* s = summ + ((size_t)(task_off[k] + fam * w + x) * NL + L) * NSUM;
            tot += s[0] + Ain * s[1] + Zin * s[2];
            Ain = fma(Ain, s[5], s[3]);
            Zin = fma(Zin, s[5], s[4]);
        }
    }
    out[(size_t)k * 2 * NL + fam * NL + L] = tot;
}

__global__ void semisep_plan(const dwba_plain_job* __restrict__ jobs, int n_jobs, int seg, int* __restrict__ off) {
    __shared__ int s_tot[1024];
    const int nt = blockDim.x;
    const int per = (n_jobs + nt - 1) / nt;
    const int k0 = threadIdx.x * per;
    const int k1 = min(n_jobs, k0 + per);
    int sum = 0;
    for (int k = k0; k < k1; ++k) sum += tasks_of(jobs[k].n_cols, seg);
    s_tot[threadIdx.x] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        int run = 0;
        for (int x = 0; x < nt; ++x) { const int v = s_tot[x]; s_tot[x] = run; run += v; }
        off[n_jobs] = run;
    }
    __syncthreads();
    int run = s_tot[threadIdx.x];
    for (int k = k0; k < k1; ++k) {
        off[k] = run;
        run += tasks_of(jobs[k].n_cols, seg);
    }
}

// ---------------------------------------------------------------------------
// Lane-serial form: one lane owns one (job, family, block of LB multipoles)
// stream and runs the recurrences over the job's whole grid, so there are no
// carries to accumulate (Ca, Cb, P) and nothing to join.  The lane carries
// A'_i = A_i + beta_i and Z'_i = Z_i + alpha_i, so each update is one FMA with
// the next point's value,
//     A'_{i+1} = q_i^L A'_i + beta_{i+1},   Z'_{i+1} = q_i^L Z'_i + alpha_{i+1},
// and the sum is taken over (alpha_i A'_i + beta_i Z'_i) / r_i with the
// diagonal coefficient d_L - 2 (alpha_i A_i + beta_i Z_i = alpha_i A'_i +
// beta_i Z'_i - 2 alpha_i beta_i): 5 FP64 operations per grid point and
// multipole; beyond the alpha rows only Z and the sum remain (3).  The
// rewrite moves at most 2 eps sum |alpha_i beta_i| / r_i of rounding, a part
// of the cancellation-free magnitude of the integral.  Lanes are
// stream-major with NBP (a power of two >= the block count) lanes per
// stream; streams [0, n_jobs) are the direct families, [n_jobs, 2 n_jobs)
// the exchange ones, so a batch whose direct families are computed
// elsewhere (alpha_dir NULL) leaves whole wavefronts idle instead of half of
// every wavefront.  Loads are grouped LGU points at a time and the next group
// is in flight while the current one is used.
constexpr int LT = 256;          // threads per workgroup of the lane-serial kernel
constexpr int LGU = 4;           // grid points per load group

// q^(blk LB): LB-th power by squaring, then blk multiplications (uniform trip count).
template <int LB>
__device__ __forceinline__ double lane_q0(double q, int blk, int nbp) {
    double qb = q;
#pragma unroll
    for (int m = 1; m < LB; m <<= 1) qb *= qb;
    double q0 = 1.0;
    for (int m = 0; m + 1 < nbp; ++m) q0 = m < blk ? q0 * qb : q0;
    return q0;
}

// Point i with alpha: uses A'_i, Z'_i, advances them with alpha/beta at i+1.
template <int LB>
__device__ __forceinline__ void lane_point_a(double ri, double r_next, double a_i, double b_i, double a_next,
                                             double b_next, int blk, int nbp, double& diag, double* Ap, double* Zp,
                                             double* acc, double& ir_cur) {
    const double iri = ir_cur;
    const double ir_next = 1.0 / r_next;
    const double q = ri * ir_next;
    const double aw = a_i * iri, bw = b_i * iri;
    diag = fma(a_i, bw, diag);
    double qL = lane_q0<LB>(q, blk, nbp);
#pragma unroll
    for (int l = 0; l < LB; ++l) {
        acc[l] = fma(aw, Ap[l], acc[l]);
        acc[l] = fma(bw, Zp[l], acc[l]);
        Ap[l] = fma(qL, Ap[l], b_next);
        Zp[l] = fma(qL, Zp[l], a_next);
        qL *= q;
    }
    ir_cur = ir_next;
}

// Point i beyond the alpha rows (Z'_i = Z_i; A no longer contributes).
template <int LB>
__device__ __forceinline__ void lane_point_b(double ri, double r_next, double b_i, int blk, int nbp, double* Zp,
                                             double* acc, double& ir_cur) {
    const double iri = ir_cur;
    const double ir_next = 1.0 / r_next;
    const double q = ri * ir_next;
    const double bw = b_i * iri;
    double qL = lane_q0<LB>(q, blk, nbp);
#pragma unroll
    for (int l = 0; l < LB; ++l) {
        acc[l] = fma(bw, Zp[l], acc[l]);
        Zp[l] *= qL;
        qL *= q;
    }
    ir_cur = ir_next;
}

// Points [0, nr) (alpha rows), LGU at a time with the next group prefetched;
// every group also holds the values at its end point (the next group's first).
template <int LB>
__device__ __forceinline__ void lane_rows(const double* __restrict__ r, const double* __restrict__ al,
                                          const double* __restrict__ be, int n, int nr, int blk, int nbp,
                                          double& diag, double* Ap, double* Zp, double* acc, double& ir_cur) {
    auto ld_r = [&](int i) { return r[min(i, n - 1)]; };
    auto ld_a = [&](int i) { return i < nr ? al[i] : 0.0; };
    auto ld_b = [&](int i) { return i < n ? be[i] : 0.0; };
    int i = 0;
    if (LGU <= nr) {
        double cr[LGU + 1], ca[LGU + 1], cb[LGU + 1];
#pragma unroll
        for (int u = 0; u <= LGU; ++u) { cr[u] = ld_r(u); ca[u] = ld_a(u); cb[u] = ld_b(u); }
#pragma unroll 1
        for (; i + LGU <= nr; i += LGU) {
            const bool more = i + 2 * LGU <= nr;
            double xr[LGU], xa[LGU], xb[LGU];
            if (more) {
#pragma unroll
                for (int u = 0; u < LGU; ++u) {
                    const int k = i + LGU + 1 + u;
                    xr[u] = ld_r(k); xa[u] = ld_a(k); xb[u] = ld_b(k);
                }
            }
#pragma unroll
            for (int u = 0; u < LGU; ++u)
                lane_point_a<LB>(cr[u], cr[u + 1], ca[u], cb[u], ca[u + 1], cb[u + 1], blk, nbp, diag, Ap, Zp, acc,
                                 ir_cur);
            if (more) {
                cr[0] = cr[LGU]; ca[0] = ca[LGU]; cb[0] = cb[LGU];
#pragma unroll
                for (int u = 0; u < LGU; ++u) { cr[u + 1] = xr[u]; ca[u + 1] = xa[u]; cb[u + 1] = xb[u]; }
            }
        }
    }
#pragma unroll 1
    for (; i < nr; ++i)
        lane_point_a<LB>(r[i], ld_r(i + 1), al[i], be[i], ld_a(i + 1), ld_b(i + 1), blk, nbp, diag, Ap, Zp, acc,
                         ir_cur);
}

// Points [nr, n): beta and the decaying Z only.
template <int LB>
__device__ __forceinline__ void lane_tail(const double* __restrict__ r, const double* __restrict__ be, int n, int i0,
                                          int blk, int nbp, double* Zp, double* acc, double& ir_cur) {
    int i = i0;
    if (i + LGU <= n) {
        double cr[LGU + 1], cb[LGU];
        cr[0] = r[i];
#pragma unroll
        for (int u = 0; u < LGU; ++u) { cr[u + 1] = r[min(i + u + 1, n - 1)]; cb[u] = be[i + u]; }
#pragma unroll 1
        for (; i + LGU <= n; i += LGU) {
            const bool more = i + 2 * LGU <= n;
            double xr[LGU], xb[LGU];
            if (more) {
#pragma unroll
                for (int u = 0; u < LGU; ++u) { xr[u] = r[min(i + LGU + u + 1, n - 1)]; xb[u] = be[i + LGU + u]; }
            }
#pragma unroll
            for (int u = 0; u < LGU; ++u) lane_point_b<LB>(cr[u], cr[u + 1], cb[u], blk, nbp, Zp, acc, ir_cur);
            if (more) {
                cr[0] = cr[LGU];
#pragma unroll
                for (int u = 0; u < LGU; ++u) { cr[u + 1] = xr[u]; cb[u] = xb[u]; }
            }
        }
    }
#pragma unroll 1
    for (; i < n; ++i) lane_point_b<LB>(r[i], r[min(i + 1, n - 1)], be[i], blk, nbp, Zp, acc, ir_cur);
}

template <int LB, int NBP>
__global__ __launch_bounds__(LT) void semisep_lane_kernel(const dwba_plain_job* __restrict__ jobs, int n_jobs,
                                                          int L_max, double* __restrict__ out) {
    const long gl = (long)blockIdx.x * LT + threadIdx.x;
    const long stream = gl / NBP;
    const int blk = (int)(gl - stream * NBP);
    if (stream >= 2L * n_jobs) return;
    const int fam = stream >= n_jobs ? 1 : 0;
    const int k = (int)(fam ? stream - n_jobs : stream);
    const int NL = L_max + 1;
    const int L0 = blk * LB;
    if (L0 >= NL) return;
    const dwba_plain_job J = jobs[k];
    if (fam == 0 && J.alpha_dir == nullptr) return;    // direct family computed elsewhere
    double* __restrict__ o = out + (size_t)k * 2 * NL + fam * NL;
    double Ap[LB], Zp[LB], acc[LB];
    double diag = 0.0;
    const int n = J.n_cols, nr = min(J.n_rows, J.n_cols);
#pragma unroll
    for (int l = 0; l < LB; ++l) acc[l] = 0.0;
    if (nr > 0 && n > 0) {
        const double* __restrict__ al = fam ? J.alpha_ex : J.alpha_dir;
        const double* __restrict__ be = fam ? J.beta_ex : J.beta_dir;
        const double a0 = al[0], b0 = be[0];
#pragma unroll
        for (int l = 0; l < LB; ++l) { Ap[l] = b0; Zp[l] = a0; }   // A_0 = Z_0 = 0
        double ir_cur = 1.0 / J.r[0];
        lane_rows<LB>(J.r, al, be, n, nr, blk, NBP, diag, Ap, Zp, acc, ir_cur);
        lane_tail<LB>(J.r, be, n, nr, blk, NBP, Zp, acc, ir_cur);
    }
    const double ldl = log(1.0 - 1e-12);
#pragma unroll
    for (int l = 0; l < LB; ++l) {
        const int L = L0 + l;
        const double dL = L == 0 ? 1.0 : exp((double)L * ldl);
        if (L < NL) o[L] = nr > 0 ? fma(dL - 2.0, diag, acc[l]) : 0.0;
    }
}

// ---------------------------------------------------------------------------
// Table-driven lane-serial form.  1 / r_i, q_i and q_i^(LB blk) depend only
// on the grid, so one pass per grid writes them (with the same operations the
// kernel above uses, hence bit-identical sums) and every lane of every pair on
// that grid reads them instead of repeating a correctly rounded division and
// the power chain at each point.  Row i of a table holds LS = 2 + NBP doubles:
//     1 / r_i,  q_i = r_i * (1 / r_{i+1}),  (q_i^LB)^b for b = 0 .. NBP-1
// (r_{i+1} clamped to the last point; the last row's q is never used).
__global__ __launch_bounds__(256) void semisep_lane_table_kernel(const dwba_lane_grid* __restrict__ grids, int lb,
                                                                 int nbp) {
    const dwba_lane_grid G = grids[blockIdx.y];
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= G.n) return;
    const int ls = 2 + nbp;
    const double ri = G.r[i];
    const double ir_next = 1.0 / G.r[min(i + 1, G.n - 1)];
    const double q = ri * ir_next;
    double qb = q;
    for (int m = 1; m < lb; m <<= 1) qb *= qb;
    double* __restrict__ t = G.tab + (size_t)i * ls;
    t[0] = 1.0 / ri;
    t[1] = q;
    double q0 = 1.0;
    for (int b = 0; b < nbp; ++b) {
        t[2 + b] = q0;
        q0 *= qb;
    }
}

// One point with alpha (the arithmetic of lane_point_a, factors from the table).
template <int LB>
__device__ __forceinline__ void lane_point_at(double iri, double q, double q0, double a_i, double b_i, double a_next,
                                              double b_next, double& diag, double* Ap, double* Zp, double* acc) {
    const double aw = a_i * iri, bw = b_i * iri;
    diag = fma(a_i, bw, diag);
    double qL = q0;
#pragma unroll
    for (int l = 0; l < LB; ++l) {
        acc[l] = fma(aw, Ap[l], acc[l]);
        acc[l] = fma(bw, Zp[l], acc[l]);
        Ap[l] = fma(qL, Ap[l], b_next);
        Zp[l] = fma(qL, Zp[l], a_next);
        qL *= q;
    }
}

// One point beyond the alpha rows (the arithmetic of lane_point_b).
template <int LB>
__device__ __forceinline__ void lane_point_bt(double iri, double q, double q0, double b_i, double* Zp, double* acc) {
    const double bw = b_i * iri;
    double qL = q0;
#pragma unroll
    for (int l = 0; l < LB; ++l) {
        acc[l] = fma(bw, Zp[l], acc[l]);
        Zp[l] *= qL;
        qL *= q;
    }
}

template <int LB, int NBP>
__global__ __launch_bounds__(LT) void semisep_lane_tab_kernel(const dwba_plain_job* __restrict__ jobs,
                                                              const double* const* __restrict__ tabs,
                                                              const int* __restrict__ out_index, int n_jobs,
                                                              int L_max, double* __restrict__ out) {
    constexpr int LS = 2 + NBP;
    const long gl = (long)blockIdx.x * LT + threadIdx.x;
    const long stream = gl / NBP;
    const int blk = (int)(gl - stream * NBP);
    if (stream >= 2L * n_jobs) return;
    const int fam = stream >= n_jobs ? 1 : 0;
    const int k = (int)(fam ? stream - n_jobs : stream);
    const int NL = L_max + 1;
    const int L0 = blk * LB;
    if (L0 >= NL) return;
    const dwba_plain_job J = jobs[k];
    if (fam == 0 && J.alpha_dir == nullptr) return;    // direct family computed elsewhere
    const int ko = out_index ? out_index[k] : k;
    double* __restrict__ o = out + (size_t)ko * 2 * NL + fam * NL;
    double Ap[LB], Zp[LB], acc[LB];
    double diag = 0.0;
    const int n = J.n_cols, nr = min(J.n_rows, J.n_cols);
#pragma unroll
    for (int l = 0; l < LB; ++l) acc[l] = 0.0;
    if (nr > 0 && n > 0) {
        const double* __restrict__ al = fam ? J.alpha_ex : J.alpha_dir;
        const double* __restrict__ be = fam ? J.beta_ex : J.beta_dir;
        const double* __restrict__ T = tabs[k];
        auto tir = [&](int i) { return T[(size_t)i * LS]; };
        auto tq = [&](int i) { return T[(size_t)i * LS + 1]; };
        auto tq0 = [&](int i) { return T[(size_t)i * LS + 2 + blk]; };
        auto ld_a = [&](int i) { return i < nr ? al[i] : 0.0; };
        auto ld_b = [&](int i) { return i < n ? be[i] : 0.0; };
        const double a0 = al[0], b0 = be[0];
#pragma unroll
        for (int l = 0; l < LB; ++l) { Ap[l] = b0; Zp[l] = a0; }   // A_0 = Z_0 = 0
        // points [0, nr): groups of LGU with the next group's loads in flight
        int i = 0;
        if (LGU <= nr) {
            double gi[LGU], gq[LGU], g0[LGU], ca[LGU + 1], cb[LGU + 1];
#pragma unroll
            for (int u = 0; u < LGU; ++u) { gi[u] = tir(u); gq[u] = tq(u); g0[u] = tq0(u); }
#pragma unroll
            for (int u = 0; u <= LGU; ++u) { ca[u] = ld_a(u); cb[u] = ld_b(u); }
#pragma unroll 1
            for (; i + LGU <= nr; i += LGU) {
                const bool more = i + 2 * LGU <= nr;
                double xi[LGU], xq[LGU], x0[LGU], xa[LGU], xb[LGU];
                if (more) {
#pragma unroll
                    for (int u = 0; u < LGU; ++u) {
                        const int kk = i + LGU + u;
                        xi[u] = tir(kk); xq[u] = tq(kk); x0[u] = tq0(kk);
                        xa[u] = ld_a(kk + 1); xb[u] = ld_b(kk + 1);
                    }
                }
#pragma unroll
                for (int u = 0; u < LGU; ++u)
                    lane_point_at<LB>(gi[u], gq[u], g0[u], ca[u], cb[u], ca[u + 1], cb[u + 1], diag, Ap, Zp, acc);
                if (more) {
                    ca[0] = ca[LGU]; cb[0] = cb[LGU];
#pragma unroll
                    for (int u = 0; u < LGU; ++u) {
                        gi[u] = xi[u]; gq[u] = xq[u]; g0[u] = x0[u];
                        ca[u + 1] = xa[u]; cb[u + 1] = xb[u];
                    }
                }
            }
        }
#pragma unroll 1
        for (; i < nr; ++i)
            lane_point_at<LB>(tir(i), tq(i), tq0(i), al[i], be[i], ld_a(i + 1), ld_b(i + 1), diag, Ap, Zp, acc);
        // points [nr, n): beta and the decaying Z only
        if (i + LGU <= n) {
            double gi[LGU], gq[LGU], g0[LGU], cb[LGU];
#pragma unroll
            for (int u = 0; u < LGU; ++u) { gi[u] = tir(i + u); gq[u] = tq(i + u); g0[u] = tq0(i + u); cb[u] = be[i + u]; }
#pragma unroll 1
            for (; i + LGU <= n; i += LGU) {
                const bool more = i + 2 * LGU <= n;
                double xi[LGU], xq[LGU], x0[LGU], xb[LGU];
                if (more) {
#pragma unroll
                    for (int u = 0; u < LGU; ++u) {
                        const int kk = i + LGU + u;
                        xi[u] = tir(kk); xq[u] = tq(kk); x0[u] = tq0(kk); xb[u] = be[kk];
                    }
                }
#pragma unroll
                for (int u = 0; u < LGU; ++u) lane_point_bt<LB>(gi[u], gq[u], g0[u], cb[u], Zp, acc);
                if (more) {
#pragma unroll
                    for (int u = 0; u < LGU; ++u) { gi[u] = xi[u]; gq[u] = xq[u]; g0[u] = x0[u]; cb[u] = xb[u]; }
                }
            }
        }
#pragma unroll 1
        for (; i < n; ++i) lane_point_bt<LB>(tir(i), tq(i), tq0(i), be[i], Zp, acc);
    }
    const double ldl = log(1.0 - 1e-12);
#pragma unroll
    for (int l = 0; l < LB; ++l) {
        const int L = L0 + l;
        const double dL = L == 0 ? 1.0 : exp((double)L * ldl);
        if (L < NL) o[L] = nr > 0 ? fma(dL - 2.0, diag, acc[l]) : 0.0;
    }
}

// ---------------------------------------------------------------------------
// Split-range form of the table-driven lane kernel: the points of a (pair,
// family, block) are divided at m between two wavefronts of one workgroup.
// The first runs [0, m) exactly as above.  The second starts at m from zero
// carries and also accumulates, with P_i = prod_{m <= j < i} q_j^L,
//     Ca = sum_{i >= m} alpha_i / r_i P_i,   Cb = sum_{i >= m} beta_i / r_i P_i,
// which is what the true carries A_m, Z_m add to its part (the recurrences are
// linear: A_i = A0_i + A_m P_i); after a barrier it reads the first wave's
// A'_m = A_m + beta_m, Z'_m = Z_m + alpha_m, sum and diagonal from LDS and
// finishes.  m balances the two waves' operation counts (5 per point and
// multipole before m, 8 after, 5 in the beta-only range), so the longest
// chain is ~0.6 of the single-lane one at 1.3x the total work.
template <int LB, bool WC>
__device__ __forceinline__ void split_point_a(double iri, double q, double q0, double a_i, double b_i,
                                              double a_next, double b_next, double& diag, double* Ap, double* Zp,
                                              double* acc, double* Ca, double* Cb, double* P) {
    const double aw = a_i * iri, bw = b_i * iri;
    diag = fma(a_i, bw, diag);
    double qL = q0;
#pragma unroll
    for (int l = 0; l < LB; ++l) {
        acc[l] = fma(aw, Ap[l], acc[l]);
        acc[l] = fma(bw, Zp[l], acc[l]);
        Ap[l] = fma(qL, Ap[l], b_next);
        Zp[l] = fma(qL, Zp[l], a_next);
        if (WC) {
            Ca[l] = fma(aw, P[l], Ca[l]);
            Cb[l] = fma(bw, P[l], Cb[l]);
            P[l] *= qL;
        }
        qL *= q;
    }
}

template <int LB>
__device__ __forceinline__ void split_point_b(double iri, double q, double q0, double b_i, double* Zp, double* acc,
                                              double* Cb, double* P) {
    const double bw = b_i * iri;
    double qL = q0;
#pragma unroll
    for (int l = 0; l < LB; ++l) {
        acc[l] = fma(bw, Zp[l], acc[l]);
        Zp[l] *= qL;
        Cb[l] = fma(bw, P[l], Cb[l]);
        P[l] *= qL;
        qL *= q;
    }
}

// alpha points [i0, i1), LGU at a time with the next group's alpha / beta in
// flight (the table rows are shared by the wavefront's pairs of one grid and
// read at use: prefetching them too spills registers)
template <int LB, int LS, bool WC>
__device__ __forceinline__ void split_rows(const double* __restrict__ T, const double* __restrict__ al,
                                           const double* __restrict__ be, int n, int nr, int i0, int i1, int blk,
                                           double& diag, double* Ap, double* Zp, double* acc, double* Ca, double* Cb,
                                           double* P) {
    auto ld_a = [&](int i) { return i < nr ? al[i] : 0.0; };
    auto ld_b = [&](int i) { return i < n ? be[i] : 0.0; };
    auto pt = [&](int i, double a_i, double b_i, double a_n, double b_n) {
        const double* t = T + (size_t)i * LS;
        split_point_a<LB, WC>(t[0], t[1], t[2 + blk], a_i, b_i, a_n, b_n, diag, Ap, Zp, acc, Ca, Cb, P);
    };
    int i = i0;
    if (i + LGU <= i1) {
        double ca[LGU + 1], cb[LGU + 1];
#pragma unroll
        for (int u = 0; u <= LGU; ++u) { ca[u] = ld_a(i + u); cb[u] = ld_b(i + u); }
#pragma unroll 1
        for (; i + LGU <= i1; i += LGU) {
            const bool more = i + 2 * LGU <= i1;
            double xa[LGU], xb[LGU];
            if (more) {
#pragma unroll
                for (int u = 0; u < LGU; ++u) { xa[u] = ld_a(i + LGU + 1 + u); xb[u] = ld_b(i + LGU + 1 + u); }
            }
#pragma unroll
            for (int u = 0; u < LGU; ++u) pt(i + u, ca[u], cb[u], ca[u + 1], cb[u + 1]);
            if (more) {
                ca[0] = ca[LGU]; cb[0] = cb[LGU];
#pragma unroll
                for (int u = 0; u < LGU; ++u) { ca[u + 1] = xa[u]; cb[u + 1] = xb[u]; }
            }
        }
    }
#pragma unroll 1
    for (; i < i1; ++i) pt(i, al[i], be[i], ld_a(i + 1), ld_b(i + 1));
}

// beta-only points [i0, n) (second wave only)
template <int LB, int LS>
__device__ __forceinline__ void split_tail(const double* __restrict__ T, const double* __restrict__ be, int n, int i0,
                                           int blk, double* Zp, double* acc, double* Cb, double* P) {
    auto pt = [&](int i, double b_i) {
        const double* t = T + (size_t)i * LS;
        split_point_b<LB>(t[0], t[1], t[2 + blk], b_i, Zp, acc, Cb, P);
    };
    int i = i0;
    if (i + LGU <= n) {
        double cb[LGU];
#pragma unroll
        for (int u = 0; u < LGU; ++u) cb[u] = be[i + u];
#pragma unroll 1
        for (; i + LGU <= n; i += LGU) {
            const bool more = i + 2 * LGU <= n;
            double xb[LGU];
            if (more) {
#pragma unroll
                for (int u = 0; u < LGU; ++u) xb[u] = be[i + LGU + u];
            }
#pragma unroll
            for (int u = 0; u < LGU; ++u) pt(i + u, cb[u]);
            if (more) {
#pragma unroll
                for (int u = 0; u < LGU; ++u) cb[u] = xb[u];
            }
        }
    }
#pragma unroll 1
    for (; i < n; ++i) pt(i, be[i]);
}

// Split point: equal operation counts for the two waves (see above).
template <int LB>
__device__ __forceinline__ int split_at(int nr, int n) {
    const long c0 = 5 * LB + 8, c1 = 8 * LB + 8, cb = 5 * LB + 8;
    const long m = (c1 * (long)nr + cb * (long)(n - nr)) / (c0 + c1);
    return (int)min((long)nr, max(0L, m));
}

template <int LB, int NBP, int OCC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(OCC)))
void semisep_lane_split_kernel(const dwba_plain_job* __restrict__ jobs,
                                                                 const double* const* __restrict__ tabs,
                                                                 const int* __restrict__ out_index, int n_jobs,
                                                                 int L_max, double* __restrict__ out) {
    constexpr int LS = 2 + NBP;
    constexpr int NX = 3 * LB + 1;                 // A', Z', sum per multipole + diagonal
    __shared__ double s_x[2][NX][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int grp = w >> 1, h = w & 1;             // wave pair grp, part h of the range
    const long gt = (long)blockIdx.x * 128 + grp * 64 + lane;
    const long stream = gt / NBP;
    const int blk = (int)(gt - stream * NBP);
    const int NL = L_max + 1;
    const int L0 = blk * LB;
    bool active = stream < 2L * n_jobs && L0 < NL;
    int fam = 0, k = 0;
    if (active) {
        fam = stream >= n_jobs ? 1 : 0;
        k = (int)(fam ? stream - n_jobs : stream);
        if (fam == 0 && jobs[k].alpha_dir == nullptr) active = false;   // direct family computed elsewhere
    }
    double Ap[LB], Zp[LB], acc[LB], Ca[LB], Cb[LB], P[LB];
    double diag = 0.0;
#pragma unroll
    for (int l = 0; l < LB; ++l) { Ap[l] = 0.0; Zp[l] = 0.0; acc[l] = 0.0; Ca[l] = 0.0; Cb[l] = 0.0; P[l] = 1.0; }
    int n = 0, nr = 0, m = 0;
    const double* __restrict__ al = nullptr;
    const double* __restrict__ be = nullptr;
    if (active) {
        const dwba_plain_job J = jobs[k];
        n = J.n_cols;
        nr = min(J.n_rows, J.n_cols);
        m = split_at<LB>(nr, n);
        al = fam ? J.alpha_ex : J.alpha_dir;
        be = fam ? J.beta_ex : J.beta_dir;
        if (nr > 0 && n > 0) {
            const double* __restrict__ T = tabs[k];
            if (h == 0) {
                const double a0 = al[0], b0 = be[0];
#pragma unroll
                for (int l = 0; l < LB; ++l) { Ap[l] = b0; Zp[l] = a0; }     // A_0 = Z_0 = 0
                split_rows<LB, LS, false>(T, al, be, n, nr, 0, m, blk, diag, Ap, Zp, acc, Ca, Cb, P);
            } else {
                const double am = m < nr ? al[m] : 0.0, bm = m < n ? be[m] : 0.0;
#pragma unroll
                for (int l = 0; l < LB; ++l) { Ap[l] = bm; Zp[l] = am; }     // zero carries at m
                split_rows<LB, LS, true>(T, al, be, n, nr, m, nr, blk, diag, Ap, Zp, acc, Ca, Cb, P);
                split_tail<LB, LS>(T, be, n, nr, blk, Zp, acc, Cb, P);
            }
        }
    }
    if (h == 0) {
#pragma unroll
        for (int l = 0; l < LB; ++l) {
            s_x[grp][l][lane] = Ap[l];
            s_x[grp][LB + l][lane] = Zp[l];
            s_x[grp][2 * LB + l][lane] = acc[l];
        }
        s_x[grp][3 * LB][lane] = diag;
    }
    __syncthreads();
    if (h == 1 && active) {
        const int ko = out_index ? out_index[k] : k;
        double* __restrict__ o = out + (size_t)ko * 2 * NL + fam * NL;
        const double am = (nr > 0 && m < nr) ? al[m] : 0.0, bm = (nr > 0 && m < n) ? be[m] : 0.0;
        const double dg = diag + s_x[grp][3 * LB][lane];
        const double ldl = log(1.0 - 1e-12);
#pragma unroll
        for (int l = 0; l < LB; ++l) {
            const double A_m = s_x[grp][l][lane] - bm;
            const double Z_m = s_x[grp][LB + l][lane] - am;
            const double tot = fma(A_m, Ca[l], fma(Z_m, Cb[l], acc[l] + s_x[grp][2 * LB + l][lane]));
            const int L = L0 + l;
            const double dL = L == 0 ? 1.0 : exp((double)L * ldl);
            if (L < NL) o[L] = nr > 0 ? fma(dL - 2.0, dg, tot) : 0.0;
        }
    }
}

// ---------------------------------------------------------------------------
// LDS-staged split form.  The 64 / NBP (job, family) streams of a wavefront
// pair run at the same grid positions: the first wavefront covers [0, m),
// the second [m, end) with the carry sums of the split kernel above, m a
// multiple of the chunk that balances the two over the streams' longest
// alpha rows and grid.  Each wavefront copies its streams' alpha / beta and
// the grid-table rows of a chunk of SSC points from HBM straight into LDS
// (global_load_lds_dword: no registers held while the copy is in flight),
// two chunks ahead of the one it computes, so the recurrences read LDS only.
// Points past a stream's own rows are fetched from a zero array, so one
// instruction stream serves every lane: a step with alpha_i = beta_i = 0 adds
// nothing to a sum, and chunks past the longest alpha rows take the
// beta-only steps.  The streams of a wavefront pair with points must share
// one grid table (the host pads each grid's jobs to whole pairs: job k with
// d_out_index[k] < 0 is not written); a pair that does not writes NaN.
//
// LDS layout of a chunk buffer (dwords): alpha copies k = 0 .. KA-1 at 64 k,
// beta at 64 (KA + k), table rows after them.  Copy k of lane l = NBP s + b
// (stream s, multipole block b) brings dword NBP k + b of stream s's chunk to
// dword 64 k + l, so point u of stream s sits at double 32 (2u / NBP) +
// (2u % NBP) / 2 past NBP s / 2: the 64 / NBP streams of a read hit distinct
// banks.
constexpr int SSC = 16;          // grid points per staged chunk

__device__ double g_stage_zeros[64];   // read in place of points past a stream's rows

typedef __attribute__((address_space(3))) void* lds_void_ptr;

// LB multipoles per lane, NBP lanes per stream, NBUF chunk buffers per
// wavefront (NBUF - 1 chunks in flight).
template <int LB_, int NBP_, int NBUF_>
struct StageShape {
    static constexpr int LB = LB_, NBP = NBP_, NBUF = NBUF_;
    static constexpr int LS = 2 + NBP;                              // table row (doubles)
    static constexpr int KA = (2 * (SSC + 1) + NBP - 1) / NBP;      // copies per array and chunk
    static constexpr int PA = KA * NBP / 2;                         // points copied per stream (>= SSC + 1)
    static constexpr int TW = SSC * 2 * LS;                         // table dwords per chunk
    static constexpr int KT = TW / 64;
    static constexpr int NI = 2 * KA + KT;                          // copies per chunk
    static constexpr int CW = 64 * NI;                              // dwords per chunk buffer
    static constexpr int OFF_B = 64 * KA, OFF_T = 128 * KA;
    static_assert(NBP % 2 == 0 && 64 % NBP == 0, "whole streams per wavefront, a double's two dwords in one copy");
    static_assert(TW % 64 == 0, "table copies fill whole wavefronts");
    static_assert(2 * NBUF * CW * 4 >= (3 * LB + 1) * 64 * 8, "join buffer fits the pair's chunk buffers");
    static_assert(NBUF == 2 || NBUF == 3, "one or two chunks in flight");
    static_assert((NBUF - 1) * NI <= 63, "the copies in flight fit the vmcnt counter");
};

// double index of point u of a stream's chunk, past the stream's own offset
__device__ __forceinline__ constexpr int stage_at(int u, int nbp) { return 32 * ((2 * u) / nbp) + ((2 * u) % nbp) / 2; }

template <int N>
__device__ __forceinline__ void stage_wait() { asm volatile("s_waitcnt vmcnt(%0)" ::"n"(N) : "memory"); }

__device__ __forceinline__ void stage_dma(const void* g, float* l) {
    __builtin_amdgcn_global_load_lds(g, (lds_void_ptr)l, 4, 0, 0);
}

// Copies of chunk [s0, s0 + SSC) into buf.  al / be: the lane's stream
// (the zero array for lanes without points, with s0l = 0); la / lbn: its
// alpha and beta lengths; clean_*: every stream's copy lies inside its rows
// (wave-uniform; otherwise each dword past them is fetched from the zeros).
template <class S>
__device__ __forceinline__ void stage_copy(float* buf, const double* al, const double* be, const double* T, int s0,
                                           int s0l, int la, int lbn, int t_rows, bool clean_a, bool clean_b,
                                           int lane) {
    constexpr int NBP = S::NBP;
    const int b = lane % NBP;
    const char* sa = reinterpret_cast<const char*>(al + s0l) + 4 * b;
    const char* sb = reinterpret_cast<const char*>(be + s0l) + 4 * b;
    if (clean_a) {
#pragma unroll
        for (int k = 0; k < S::KA; ++k) stage_dma(sa + 4 * NBP * k, buf + 64 * k);
    } else {
#pragma unroll
        for (int k = 0; k < S::KA; ++k) {
            const bool in = s0l + (NBP * k + b) / 2 < la;
            stage_dma(in ? (const void*)(sa + 4 * NBP * k) : (const void*)g_stage_zeros, buf + 64 * k);
        }
    }
    if (clean_b) {
#pragma unroll
        for (int k = 0; k < S::KA; ++k) stage_dma(sb + 4 * NBP * k, buf + S::OFF_B + 64 * k);
    } else {
#pragma unroll
        for (int k = 0; k < S::KA; ++k) {
            const bool in = s0l + (NBP * k + b) / 2 < lbn;
            stage_dma(in ? (const void*)(sb + 4 * NBP * k) : (const void*)g_stage_zeros, buf + S::OFF_B + 64 * k);
        }
    }
    // table rows s0 .. s0 + SSC - 1, row-major, rows past the grid clamped to its last
    const float* tf = reinterpret_cast<const float*>(T);
    const long d0 = (long)s0 * 2 * S::LS + lane;
    if (s0 + SSC <= t_rows) {
#pragma unroll
        for (int j = 0; j < S::KT; ++j) stage_dma(tf + d0 + 64 * j, buf + S::OFF_T + 64 * j);
    } else {
        const long dmax = (long)t_rows * 2 * S::LS - 1;
#pragma unroll
        for (int j = 0; j < S::KT; ++j) stage_dma(tf + min(d0 + 64 * j, dmax), buf + S::OFF_T + 64 * j);
    }
}

// One chunk of points with alpha (the arithmetic of split_point_a).
template <class S, bool WC>
__device__ __forceinline__ void stage_rows(const double* pa, const double* pb, const double* pt, const double* pq,
                                           double& diag, double* Ap, double* Zp, double* acc, double* Ca, double* Cb,
                                           double* P) {
    double a_u = pa[0], b_u = pb[0];
#pragma unroll
    for (int u = 0; u < SSC; ++u) {
        const double a_n = pa[stage_at(u + 1, S::NBP)];
        const double b_n = pb[stage_at(u + 1, S::NBP)];
        split_point_a<S::LB, WC>(pt[S::LS * u], pt[S::LS * u + 1], pq[S::LS * u], a_u, b_u, a_n, b_n, diag, Ap, Zp,
                                 acc, Ca, Cb, P);
        a_u = a_n;
        b_u = b_n;
    }
}

// One chunk of beta-only points (lane_point_bt; with carries split_point_b).
template <class S, bool WC>
__device__ __forceinline__ void stage_tail(const double* pb, const double* pt, const double* pq, double* Zp,
                                           double* acc, double* Cb, double* P) {
#pragma unroll
    for (int u = 0; u < SSC; ++u) {
        const double b_u = pb[stage_at(u, S::NBP)];
        if (WC) split_point_b<S::LB>(pt[S::LS * u], pt[S::LS * u + 1], pq[S::LS * u], b_u, Zp, acc, Cb, P);
        else lane_point_bt<S::LB>(pt[S::LS * u], pt[S::LS * u + 1], pq[S::LS * u], b_u, Zp, acc);
    }
}

// Chunks [c0, c1) of one wavefront, alpha steps below chunk ca.
template <class S, bool WC>
__device__ __forceinline__ void stage_run(float* mine, int c0, int c1, int ca, const double* al, const double* be,
                                          const double* T, bool has, int la, int lbn, int la_min, int lb_min,
                                          int t_rows, int lane, double& diag, double* Ap, double* Zp, double* acc,
                                          double* Ca, double* Cb, double* P) {
    constexpr int NBP = S::NBP, NBUF = S::NBUF;
    auto copy = [&](int c) {
        const int s0 = c * SSC;
        stage_copy<S>(mine + (c % NBUF) * S::CW, al, be, T, s0, has ? s0 : 0, la, lbn, t_rows,
                      s0 + S::PA <= la_min, s0 + S::PA <= lb_min, lane);
    };
    if (c0 < c1) copy(c0);
    if (NBUF == 3 && c0 + 1 < c1) copy(c0 + 1);
#pragma unroll 1
    for (int c = c0; c < c1; ++c) {
        if (c + NBUF - 1 < c1) {
            copy(c + NBUF - 1);
            stage_wait<(NBUF - 1) * S::NI>();
        } else if (NBUF == 3 && c + 1 < c1) {
            stage_wait<S::NI>();
        } else {
            stage_wait<0>();
        }
        const float* buf = mine + (c % NBUF) * S::CW;
        const double* pa = reinterpret_cast<const double*>(buf) + (lane / NBP) * (NBP / 2);
        const double* pb = pa + S::OFF_B / 2;
        const double* pt = reinterpret_cast<const double*>(buf + S::OFF_T);
        const double* pq = pt + 2 + lane % NBP;
        if (c < ca) stage_rows<S, WC>(pa, pb, pt, pq, diag, Ap, Zp, acc, Ca, Cb, P);
        else stage_tail<S, WC>(pb, pt, pq, Zp, acc, Cb, P);
    }
    stage_wait<0>();
}

__device__ __forceinline__ int wave_max(int x) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) x = max(x, __shfl_xor(x, d, 64));
    return x;
}
__device__ __forceinline__ int wave_min(int x) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) x = min(x, __shfl_xor(x, d, 64));
    return x;
}

template <int LB, int NBP, int NBUF, int OCC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(OCC)))
void semisep_lane_stage_kernel(const dwba_plain_job* __restrict__ jobs, const double* const* __restrict__ tabs,
                               const int* __restrict__ out_index, int n_jobs, int L_max, double* __restrict__ out) {
    using S = StageShape<LB, NBP, NBUF>;
    __shared__ __attribute__((aligned(16))) float s_stage[4][NBUF][S::CW];
    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int grp = w >> 1, h = w & 1;              // wave pair grp, part h of the range
    const long gt = (long)blockIdx.x * 128 + grp * 64 + lane;
    const long stream = gt / NBP;
    const int blk = lane % NBP;
    const int NL = L_max + 1;
    const int L0 = blk * LB;
    bool ok = stream < 2L * n_jobs;
    int fam = 0, k = 0, n = 0, nr = 0;
    const double* al = g_stage_zeros;
    const double* be = g_stage_zeros;
    const double* T = nullptr;
    if (ok) {
        fam = stream >= n_jobs ? 1 : 0;
        k = (int)(fam ? stream - n_jobs : stream);
        const dwba_plain_job J = jobs[k];
        const double* a = fam ? J.alpha_ex : J.alpha_dir;
        if (a == nullptr) {
            ok = false;                                // direct family computed elsewhere
        } else if (J.n_cols > 0) {
            al = a;
            be = fam ? J.beta_ex : J.beta_dir;
            n = J.n_cols;
            nr = max(0, min(J.n_rows, J.n_cols));
            T = tabs[k];
        }
    }
    const bool has = n > 0;                             // a stream with points (its lanes copy and compute)
    // pair-uniform extents: longest grid and alpha rows, shortest of each for the copies' bounds
    const int e_w = __builtin_amdgcn_readfirstlane(wave_max(n)), r_w = __builtin_amdgcn_readfirstlane(wave_max(nr));
    const int la_min = __builtin_amdgcn_readfirstlane(wave_min(has ? nr : INT_MAX));
    const int lb_min = __builtin_amdgcn_readfirstlane(wave_min(has ? n : INT_MAX));
    const unsigned long long hm = __ballot(has);
    const double* T0 = nullptr;
    bool mixed = false;
    if (hm) {
        const int first = __ffsll((long long)hm) - 1;
        const unsigned long long tp = reinterpret_cast<unsigned long long>(T);
        const unsigned lo = __builtin_amdgcn_readlane((unsigned)tp, first);
        const unsigned hi = __builtin_amdgcn_readlane((unsigned)(tp >> 32), first);
        T0 = reinterpret_cast<const double*>(((unsigned long long)hi << 32) | lo);
        mixed = __ballot(has && T != T0) != 0;
    }
    int m_w = 0, c_end = 0, c_alpha = 0;
    if (e_w > 0 && !mixed) {
        c_end = (e_w + SSC - 1) / SSC;
        c_alpha = (r_w + SSC - 1) / SSC;
        // balance 5 (first) against 8 (second) operations per point and multipole on the alpha rows, 5
        // against 3 past them (beta-only steps, with and without the carry sums)
        const long c0 = 5 * LB + 8, c1 = 8 * LB + 8, cb = 5 * LB + 8, c0b = 3 * LB + 8;
        long m = (c1 * r_w + cb * (long)(e_w - r_w)) / (c0 + c1);
        if (m > r_w) m = ((c0b - c0) * (long)r_w + cb * (long)e_w) / (c0b + cb);
        m_w = (int)min((long)c_end, max(0L, (m + SSC / 2) / SSC)) * SSC;
    }
    const int i0 = h ? m_w : 0;
    const double a0 = (has && i0 < nr) ? al[i0] : 0.0, b0 = (has && i0 < n) ? be[i0] : 0.0;
    double Ap[LB], Zp[LB], acc[LB], Ca[LB], Cb[LB], P[LB];
    double diag = 0.0;
#pragma unroll
    for (int l = 0; l < LB; ++l) { Ap[l] = b0; Zp[l] = a0; acc[l] = 0.0; Ca[l] = 0.0; Cb[l] = 0.0; P[l] = 1.0; }
    const int la = has ? nr : INT_MAX, lbn = has ? n : INT_MAX;
    float* mine = &s_stage[w][0][0];
    if (h == 0)
        stage_run<S, false>(mine, 0, m_w / SSC, c_alpha, al, be, T0, has, la, lbn, la_min, lb_min, e_w, lane, diag,
                            Ap, Zp, acc, Ca, Cb, P);
    else
        stage_run<S, true>(mine, m_w / SSC, c_end, c_alpha, al, be, T0, has, la, lbn, la_min, lb_min, e_w, lane,
                           diag, Ap, Zp, acc, Ca, Cb, P);
    __syncthreads();                                    // every copy landed and read: the buffers are free
    double* s_x = reinterpret_cast<double*>(&s_stage[2 * grp][0][0]);   // [3 LB + 1][64]
    if (h == 0) {
#pragma unroll
        for (int l = 0; l < LB; ++l) {
            s_x[l * 64 + lane] = Ap[l];
            s_x[(LB + l) * 64 + lane] = Zp[l];
            s_x[(2 * LB + l) * 64 + lane] = acc[l];
        }
        s_x[3 * LB * 64 + lane] = diag;
    }
    __syncthreads();
    if (h == 1 && ok && L0 < NL) {
        const int ko = out_index ? out_index[k] : k;
        if (ko >= 0) {
            double* __restrict__ o = out + (size_t)ko * 2 * NL + fam * NL;
            const double dg = diag + s_x[3 * LB * 64 + lane];
            const double ldl = log(1.0 - 1e-12);
#pragma unroll
            for (int l = 0; l < LB; ++l) {
                const double A_m = s_x[l * 64 + lane] - b0;
                const double Z_m = s_x[(LB + l) * 64 + lane] - a0;
                const double tot = fma(A_m, Ca[l], fma(Z_m, Cb[l], acc[l] + s_x[(2 * LB + l) * 64 + lane]));
                const int L = L0 + l;
                const double dL = L == 0 ? 1.0 : exp((double)L * ldl);
                if (L < NL) o[L] = mixed ? __builtin_nan("") : (nr > 0 ? fma(dL - 2.0, dg, tot) : 0.0);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Shared-beta form (the direct family of an excitation scan, where every
// pair of an energy has the same beta = u_f u_i w): t_L(i) = sum_j K_L(i, j) beta_j
// for i < rows, all L, once per energy, in O(N) per L:
//     t_L(i) = (A_i + d_L beta_i) / r_i + B_i,
//     A_{i+1} = q_i^L (A_i + beta_i),              A_0 = 0          (j < i)
//     B_{i-1} = q_{i-1}^L (B_i + beta_i / r_i),    B_{N-1} = 0      (j > i)
constexpr int GT = 256;

// A workgroup per (energy, L) sweeps the grid in chunks of GT * GPT points
// (GPT consecutive points per thread, below).  Point i's step of the
// forward recurrence is the affine map x -> q_i^L x + q_i^L beta_i (of the
// backward one x -> q_{i-1}^L x + q_{i-1}^L beta_i / r_i); a chunk composes
// them with an in-wave shuffle scan and a scan of the four wave totals in
// LDS, applies the carry from the previous chunk, and writes t (forward
// sweep) or adds to it (backward sweep) with coalesced accesses.
struct Aff { double m, v; };                     // x -> m x + v

__device__ __forceinline__ Aff aff_then(Aff first, Aff second) {   // second after first
    return Aff{second.m * first.m, fma(second.m, first.v, second.v)};
}

// Inclusive scan over the workgroup in thread order (rev: reversed order);
// returns this thread's exclusive prefix, and the whole chunk in *tot.
__device__ __forceinline__ Aff block_scan_excl(Aff x, bool rev, double* s_m, double* s_v, Aff* tot) {
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int pos = rev ? 63 - lane : lane;                  // position in the scan order within the wave
    Aff inc = x;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int src = rev ? lane + d : lane - d;
        const double pm = __shfl(inc.m, src < 0 || src > 63 ? lane : src, 64);
        const double pv = __shfl(inc.v, src < 0 || src > 63 ? lane : src, 64);
        if (pos >= d) inc = aff_then(Aff{pm, pv}, inc);
    }
    const int last = rev ? 0 : 63;                            // lane holding the wave's total
    if (lane == last) { s_m[w] = inc.m; s_v[w] = inc.v; }
    __syncthreads();
    constexpr int NW = GT / 64;
    Aff pre{1.0, 0.0};                                        // earlier waves in scan order
    for (int q = 0; q < NW; ++q) {
        const int wq = rev ? NW - 1 - q : q;
        if (wq == w) break;
        pre = aff_then(pre, Aff{s_m[wq], s_v[wq]});
    }
    Aff all = Aff{1.0, 0.0};
    for (int q = 0; q < NW; ++q) {
        const int wq = rev ? NW - 1 - q : q;
        all = aff_then(all, Aff{s_m[wq], s_v[wq]});
    }
    *tot = all;
    // exclusive = the previous element's inclusive value (pre for the wave's first)
    const int src = rev ? lane + 1 : lane - 1;
    const double em = __shfl(inc.m, src < 0 || src > 63 ? lane : src, 64);
    const double ev = __shfl(inc.v, src < 0 || src > 63 ? lane : src, 64);
    const Aff ex = pos == 0 ? Aff{1.0, 0.0} : Aff{em, ev};
    __syncthreads();                                          // s_m / s_v are reused by the next call
    return aff_then(pre, ex);
}

// Each thread takes GPT consecutive points of a chunk of GT * GPT: it
// composes their affine steps serially, the workgroup scans the GT
// compositions, and the thread replays its points from its exclusive prefix
// -- one block scan (two barriers) per GT * GPT points instead of per GT.
constexpr int GPT = 4;

__global__ __launch_bounds__(GT) void semisep_gemv_scan_kernel(const dwba_dgemv_energy* __restrict__ en) {
    __shared__ double s_m[GT / 64], s_v[GT / 64];
    const dwba_dgemv_energy E = en[blockIdx.x];
    const int L = blockIdx.y;
    if (L > E.L_max) return;
    const int n = E.N, rows = E.rows, tid = threadIdx.x;
    const double* __restrict__ r = E.r;
    const double* __restrict__ be = E.beta;
    double* __restrict__ t = E.t + (size_t)L * E.ld;
    const double dL = L == 0 ? 1.0 : exp((double)L * log(1.0 - 1e-12));
    const int top = min(rows, n);                             // t is needed for i < rows only
    constexpr int CH = GT * GPT;
    // forward: A_{i+1} = q_i^L (A_i + beta_i), A_0 = 0; t_i = (A_i + d_L beta_i) / r_i
    double carry = 0.0;
    for (int c0 = 0; c0 < top; c0 += CH) {
        const int i0 = c0 + tid * GPT;
        double rr[GPT], bb[GPT], qq[GPT];
        Aff f{1.0, 0.0};                                      // the thread's points, first to last
#pragma unroll
        for (int k = 0; k < GPT; ++k) {
            const int i = i0 + k;
            rr[k] = 1.0; bb[k] = 0.0; qq[k] = 1.0;
            if (i < n) {
                rr[k] = r[i];
                bb[k] = be[i];
                const double q = i + 1 < n ? rr[k] / r[i + 1] : 1.0;
                qq[k] = L == 0 ? 1.0 : pow_int_rt(q, L);
                f = aff_then(f, Aff{qq[k], qq[k] * bb[k]});
            }
        }
        Aff tot;
        const Aff ex = block_scan_excl(f, false, s_m, s_v, &tot);
        double A = fma(ex.m, carry, ex.v);
#pragma unroll
        for (int k = 0; k < GPT; ++k) {
            const int i = i0 + k;
            if (i < top) t[i] = (A + dL * bb[k]) / rr[k];
            A = qq[k] * (A + bb[k]);
        }
        carry = fma(tot.m, carry, tot.v);
    }
    // backward: B_{i-1} = q_{i-1}^L (B_i + beta_i / r_i), B_{n-1} = 0; t_i += B_i.
    // Chunks from the top of the grid down, each thread's points last to first;
    // points >= top only feed the carry.
    carry = 0.0;
    const int nch = (n + CH - 1) / CH;
    for (int c = nch - 1; c >= 0; --c) {
        const int i0 = c * CH + tid * GPT;
        double qq[GPT], ww[GPT];
        Aff g{1.0, 0.0};                                      // the thread's points, last to first
#pragma unroll
        for (int k = GPT - 1; k >= 0; --k) {
            const int i = i0 + k;
            qq[k] = 1.0; ww[k] = 0.0;
            if (i < n) {
                const double ri = r[i];
                const double q = i > 0 ? r[i - 1] / ri : 1.0;
                qq[k] = L == 0 ? 1.0 : pow_int_rt(q, L);
                ww[k] = be[i] / ri;
                g = aff_then(g, Aff{qq[k], qq[k] * ww[k]});
            }
        }
        Aff tot;
        const Aff ex = block_scan_excl(g, true, s_m, s_v, &tot);
        double B = fma(ex.m, carry, ex.v);
#pragma unroll
        for (int k = GPT - 1; k >= 0; --k) {
            const int i = i0 + k;
            if (i < top) t[i] += B;
            B = qq[k] * (B + ww[k]);
        }
        carry = fma(tot.m, carry, tot.v);
    }
}

// ---------------------------------------------------------------------------
// O(N) gradients of sum_{p,L} g[p][L] I_L(p), I_L(p) = (w rho_a[p])^T K_L (w rho_b[p]):
//     grad_a[p] = w (sum_L g[p][L] K_L (w rho_b[p])),  grad_b[p] = w (sum_L g[p][L] K_L (w rho_a[p]))
// (K_L symmetric).  Row k < P of the work is K_L (w rho_b[k]), row P + k is
// K_L (w rho_a[k]); each is one shared-beta GEMV of semisep_gemv_scan_kernel.
__global__ __launch_bounds__(GT) void backward_setup_kernel(const double* __restrict__ r, const double* __restrict__ w,
                                                            int n, int P, int ld_in, const double* __restrict__ rho_a,
                                                            const double* __restrict__ rho_b, int L_max, int ld_t,
                                                            double* __restrict__ xw, double* __restrict__ t,
                                                            dwba_dgemv_energy* __restrict__ desc) {
    const int k = blockIdx.y;                                 // 0 .. 2P-1
    const double* __restrict__ src = k < P ? rho_b + (size_t)k * ld_in : rho_a + (size_t)(k - P) * ld_in;
    double* __restrict__ dst = xw + (size_t)k * ld_t;
    for (int i = blockIdx.x * GT + threadIdx.x; i < n; i += gridDim.x * GT) dst[i] = w[i] * src[i];
    if (blockIdx.x == 0 && threadIdx.x == 0)
        desc[k] = dwba_dgemv_energy{r, dst, t + (size_t)k * (L_max + 1) * ld_t, n, n, ld_t, L_max};
}

__global__ __launch_bounds__(GT) void backward_combine_kernel(const double* __restrict__ w, int n, int P, int ld_out,
                                                              const double* __restrict__ g, int L_max, int ld_t,
                                                              const double* __restrict__ t, double* __restrict__ ga,
                                                              double* __restrict__ gb) {
    const int k = blockIdx.y;
    const int p = k < P ? k : k - P;
    double* __restrict__ out = k < P ? ga : gb;
    if (out == nullptr) return;
    out += (size_t)p * ld_out;
    const double* __restrict__ tk = t + (size_t)k * (L_max + 1) * ld_t;
    const double* __restrict__ gp = g + (size_t)p * (L_max + 1);
    for (int i = blockIdx.x * GT + threadIdx.x; i < ld_out; i += gridDim.x * GT) {
        double v = 0.0;
        if (i < n) {
            double s = 0.0;
            for (int L = 0; L <= L_max; ++L) s = fma(gp[L], tk[(size_t)L * ld_t + i], s);
            v = w[i] * s;
        }
        out[i] = v;
    }
}

struct BackwardLayout {
    int ld_t;
    size_t desc_off, xw_off, t_off, total;
};
BackwardLayout backward_layout(int n, int P, int L_max) {
    BackwardLayout b;
    b.ld_t = (n + 31) / 32 * 32;
    b.desc_off = 0;
    b.xw_off = align_up(sizeof(dwba_dgemv_energy) * 2 * (size_t)P);
    b.t_off = b.xw_off + align_up(sizeof(double) * 2 * (size_t)P * b.ld_t);
    b.total = b.t_off + align_up(sizeof(double) * 2 * (size_t)P * (L_max + 1) * b.ld_t);
    return b;
}

int host_tasks(int32_t n_jobs, const int32_t* h_n_cols, int seg) {
    int64_t tot = 0;
    for (int k = 0; k < n_jobs; ++k) {
        const int n = h_n_cols[k];
        if (n > 0) tot += 2 * ((n + 64 * seg - 1) / (64 * seg));
    }
    return tot > INT32_MAX ? -1 : (int)tot;
}

// Lane range length: enough tasks to give every SIMD several waves, no
// shorter than needed to keep tasks below 2^31.
int pick_seg(int32_t n_jobs, const int32_t* h_n_cols, int seg_hint) {
    if (seg_hint > 0) return std::min(seg_hint, 1 << 16);
    int64_t elems = 0;
    for (int k = 0; k < n_jobs; ++k) elems += 2 * (int64_t)std::max(0, h_n_cols[k]);
    const int64_t target_waves = (int64_t)std::max(1, device_cu_count()) * 16;
    int64_t seg = elems / (64 * target_waves);
    return (int)std::min<int64_t>(std::max<int64_t>(seg, 1), 1 << 16);
}

struct SemisepLayout {
    int seg, n_tasks;
    size_t off_bytes, summ_bytes, total;
};

bool semisep_layout(int32_t n_jobs, const int32_t* h_n_rows, const int32_t* h_n_cols, int32_t L_max, int seg_hint,
                    SemisepLayout* lay) {
    if (n_jobs <= 0 || !h_n_rows || !h_n_cols || L_max < 0 || L_max > 63) return false;
    for (int k = 0; k < n_jobs; ++k)
        if (h_n_cols[k] < 0 || h_n_rows[k] < 0 || h_n_rows[k] > h_n_cols[k]) return false;
    lay->seg = pick_seg(n_jobs, h_n_cols, seg_hint);
    lay->n_tasks = host_tasks(n_jobs, h_n_cols, lay->seg);
    if (lay->n_tasks < 0) return false;
    lay->off_bytes = align_up(sizeof(int) * (size_t)(n_jobs + 1));
    lay->summ_bytes = align_up(sizeof(double) * (size_t)std::max(1, lay->n_tasks) * (L_max + 1) * NSUM);
    lay->total = lay->off_bytes + lay->summ_bytes;
    return true;
}

}  // namespace
}  // namespace dwba

using namespace dwba;

extern "C" {

size_t dwba_il_semisep_f64_workspace_bytes(int32_t n_jobs, const int32_t* h_n_rows, const int32_t* h_n_cols,
                                           int32_t L_max, int32_t seg) {
    SemisepLayout lay;
    if (!semisep_layout(n_jobs, h_n_rows, h_n_cols, L_max, seg, &lay)) return 0;
    return lay.total;
}

int32_t dwba_il_semisep_f64(int32_t n_jobs, const dwba_plain_job* d_jobs, const int32_t* h_n_rows,
                            const int32_t* h_n_cols, int32_t L_max, int32_t seg, int32_t lb, double* d_out,
                            void* d_workspace, size_t workspace_bytes, void* stream) {
    DWBA_REQUIRE(initialised(), "dwba_il_semisep_f64: call dwba_init first");
    DWBA_REQUIRE(d_jobs && d_out, "dwba_il_semisep_f64: null pointer");
    if (lb == 0) lb = 8;
    DWBA_REQUIRE(lb == 4 || lb == 8, "dwba_il_semisep_f64: lb must be 0, 4 or 8 (got %d)", lb);
    SemisepLayout lay;
    DWBA_REQUIRE(semisep_layout(n_jobs, h_n_rows, h_n_cols, L_max, seg, &lay),
                 "dwba_il_semisep_f64: invalid jobs (n_jobs=%d, L_max=%d; need 0 <= n_rows <= n_cols)", n_jobs, L_max);
    if (workspace_bytes < lay.total)
        return set_error(DWBA_ERR_WORKSPACE, "dwba_il_semisep_f64: workspace %zu < %zu bytes", workspace_bytes,
                         lay.total);
    DWBA_REQUIRE(d_workspace, "dwba_il_semisep_f64: null workspace");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    char* ws = static_cast<char*>(d_workspace);
    int* d_off = reinterpret_cast<int*>(ws);
    double* d_summ = reinterpret_cast<double*>(ws + lay.off_bytes);
    hipLaunchKernelGGL(semisep_plan, dim3(1), dim3(256), 0, s, d_jobs, n_jobs, lay.seg, d_off);
    DWBA_LAUNCH_CHECK("semisep_plan");
    if (lay.n_tasks > 0) {
        const int blocks = (lay.n_tasks + ST / 64 - 1) / (ST / 64);
        const int nb = (L_max + lb) / lb;     // L-blocks of lb multipoles
        const int key = lb * 100 + nb;
        switch (key) {
#define C(LBV, NBV) case LBV * 100 + NBV: hipLaunchKernelGGL((semisep_kernel<LBV, NBV>), dim3(blocks), dim3(ST), 0, s, \
                                                   d_jobs, d_off, n_jobs, lay.seg, L_max, d_summ); break;
            C(8, 1) C(8, 2) C(8, 3) C(8, 4) C(8, 5) C(8, 6) C(8, 7) C(8, 8)
            C(4, 1) C(4, 2) C(4, 3) C(4, 4) C(4, 5) C(4, 6) C(4, 7) C(4, 8)
            C(4, 9) C(4, 10) C(4, 11) C(4, 12) C(4, 13) C(4, 14) C(4, 15) C(4, 16)
#undef C
            default: return set_error(DWBA_ERR_ARG, "dwba_il_semisep_f64: L_max %d", L_max);
        }
        DWBA_LAUNCH_CHECK("semisep_kernel");
    }
    hipLaunchKernelGGL(semisep_join, dim3(n_jobs, 2), dim3(64), 0, s, d_jobs, d_off, lay.seg, L_max, d_summ, d_out);
    DWBA_LAUNCH_CHECK("semisep_join");
    return DWBA_OK;
}

int32_t dwba_il_semisep_lanes_f64(int32_t n_jobs, const dwba_plain_job* d_jobs, int32_t L_max, int32_t lb,
                                  double* d_out, void* stream) {
    DWBA_REQUIRE(initialised(), "dwba_il_semisep_lanes_f64: call dwba_init first");
    DWBA_REQUIRE(n_jobs > 0 && d_jobs && d_out && L_max >= 0 && L_max <= 63,
                 "dwba_il_semisep_lanes_f64: bad arguments (n_jobs=%d, L_max=%d)", n_jobs, L_max);
    if (lb == 0) lb = 8;
    DWBA_REQUIRE(lb == 4 || lb == 8, "dwba_il_semisep_lanes_f64: lb must be 0, 4 or 8 (got %d)", lb);
    const int nb = (L_max + lb) / lb;
    int nbp = 1;
    while (nbp < nb) nbp <<= 1;                      // 1 .. 16
    const long lanes = 2L * n_jobs * nbp;
    const long blocks = (lanes + LT - 1) / LT;
    DWBA_REQUIRE(blocks <= INT32_MAX, "dwba_il_semisep_lanes_f64: too many jobs (%d)", n_jobs);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    switch (lb * 100 + nbp) {
#define C(B, P) case B * 100 + P: hipLaunchKernelGGL((semisep_lane_kernel<B, P>), dim3((unsigned)blocks), dim3(LT), 0, \
                                                     s, d_jobs, n_jobs, L_max, d_out); break;
        C(8, 1) C(8, 2) C(8, 4) C(8, 8) C(4, 1) C(4, 2) C(4, 4) C(4, 8) C(4, 16)
#undef C
        default: return set_error(DWBA_ERR_ARG, "dwba_il_semisep_lanes_f64: L_max %d", L_max);
    }
    DWBA_LAUNCH_CHECK("semisep_lane_kernel");
    return DWBA_OK;
}

static int lane_nbp(int32_t L_max, int32_t lb) {
    const int nb = (L_max + lb) / lb;
    int nbp = 1;
    while (nbp < nb) nbp <<= 1;                      // 1 .. 16
    return nbp;
}

int32_t dwba_semisep_lane_table_stride(int32_t L_max, int32_t lb) {
    if (lb == 0) lb = 8;
    if ((lb != 4 && lb != 8) || L_max < 0 || L_max > 63) return -1;
    return 2 + lane_nbp(L_max, lb);
}

int32_t dwba_semisep_lane_tables_f64(int32_t n_grids, const dwba_lane_grid* d_grids, int32_t max_n, int32_t L_max,
                                     int32_t lb, void* stream) {
    DWBA_REQUIRE(initialised(), "dwba_semisep_lane_tables_f64: call dwba_init first");
    if (lb == 0) lb = 8;
    DWBA_REQUIRE(n_grids > 0 && n_grids <= 65535 && d_grids && max_n > 0 && L_max >= 0 && L_max <= 63 &&
                     (lb == 4 || lb == 8),
                 "dwba_semisep_lane_tables_f64: bad arguments (n_grids=%d, max_n=%d, L_max=%d, lb=%d)", n_grids,
                 max_n, L_max, lb);
    hipLaunchKernelGGL(semisep_lane_table_kernel, dim3((max_n + 255) / 256, n_grids), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), d_grids, lb, lane_nbp(L_max, lb));
    DWBA_LAUNCH_CHECK("semisep_lane_table_kernel");
    return DWBA_OK;
}

int32_t dwba_il_semisep_lanes_tab_f64(int32_t n_jobs, const dwba_plain_job* d_jobs, const double* const* d_tabs,
                                      const int32_t* d_out_index, int32_t L_max, int32_t lb, double* d_out,
                                      void* stream) {
    DWBA_REQUIRE(initialised(), "dwba_il_semisep_lanes_tab_f64: call dwba_init first");
    DWBA_REQUIRE(n_jobs > 0 && d_jobs && d_tabs && d_out && L_max >= 0 && L_max <= 63,
                 "dwba_il_semisep_lanes_tab_f64: bad arguments (n_jobs=%d, L_max=%d)", n_jobs, L_max);
    if (lb == 0) lb = 8;
    DWBA_REQUIRE(lb == 4 || lb == 8, "dwba_il_semisep_lanes_tab_f64: lb must be 0, 4 or 8 (got %d)", lb);
    const int nbp = lane_nbp(L_max, lb);
    const long lanes = 2L * n_jobs * nbp;
    const long blocks = (lanes + LT - 1) / LT;
    DWBA_REQUIRE(blocks <= INT32_MAX, "dwba_il_semisep_lanes_tab_f64: too many jobs (%d)", n_jobs);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    switch (lb * 100 + nbp) {
#define C(B, P) case B * 100 + P: hipLaunchKernelGGL((semisep_lane_tab_kernel<B, P>), dim3((unsigned)blocks), \
                                                     dim3(LT), 0, s, d_jobs, d_tabs, d_out_index, n_jobs, L_max, \
                                                     d_out); break;
        C(8, 1) C(8, 2) C(8, 4) C(8, 8) C(4, 1) C(4, 2) C(4, 4) C(4, 8) C(4, 16)
#undef C
        default: return set_error(DWBA_ERR_ARG, "dwba_il_semisep_lanes_tab_f64: L_max %d", L_max);
    }
    DWBA_LAUNCH_CHECK("semisep_lane_tab_kernel");
    return DWBA_OK;
}

int32_t dwba_il_semisep_lanes_split_f64(int32_t n_jobs, const dwba_plain_job* d_jobs, const double* const* d_tabs,
                                        const int32_t* d_out_index, int32_t L_max, int32_t lb,
                                        int32_t waves_per_simd, double* d_out, void* stream) {
    DWBA_REQUIRE(initialised(), "dwba_il_semisep_lanes_split_f64: call dwba_init first");
    DWBA_REQUIRE(n_jobs > 0 && d_jobs && d_tabs && d_out && L_max >= 0 && L_max <= 63,
                 "dwba_il_semisep_lanes_split_f64: bad arguments (n_jobs=%d, L_max=%d)", n_jobs, L_max);
    if (lb == 0) lb = 8;
    DWBA_REQUIRE(lb == 4 || lb == 8, "dwba_il_semisep_lanes_split_f64: lb must be 0, 4 or 8 (got %d)", lb);
    const int nbp = lane_nbp(L_max, lb);
    const long tasks = 2L * n_jobs * nbp;            // (job, family, block); two wavefronts per 64 of them
    const long blocks = (tasks + 127) / 128;
    DWBA_REQUIRE(blocks <= INT32_MAX, "dwba_il_semisep_lanes_split_f64: too many jobs (%d)", n_jobs);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (waves_per_simd == 0) waves_per_simd = 2;
    DWBA_REQUIRE(waves_per_simd == 2 || waves_per_simd == 3,
                 "dwba_il_semisep_lanes_split_f64: waves_per_simd must be 0, 2 or 3 (got %d)", waves_per_simd);
    switch (waves_per_simd * 10000 + lb * 100 + nbp) {
#define C(O, B, P) case O * 10000 + B * 100 + P: \
        hipLaunchKernelGGL((semisep_lane_split_kernel<B, P, O>), dim3((unsigned)blocks), dim3(256), 0, s, d_jobs, \
                           d_tabs, d_out_index, n_jobs, L_max, d_out); break;
        C(2, 8, 1) C(2, 8, 2) C(2, 8, 4) C(2, 8, 8) C(2, 4, 1) C(2, 4, 2) C(2, 4, 4) C(2, 4, 8) C(2, 4, 16)
        C(3, 8, 1) C(3, 8, 2) C(3, 8, 4) C(3, 8, 8) C(3, 4, 1) C(3, 4, 2) C(3, 4, 4) C(3, 4, 8) C(3, 4, 16)
#undef C
        default: return set_error(DWBA_ERR_ARG, "dwba_il_semisep_lanes_split_f64: L_max %d", L_max);
    }
    DWBA_LAUNCH_CHECK("semisep_lane_split_kernel");
    return DWBA_OK;
}

static int stage_lb(int32_t lb) { return lb == 0 ? 4 : lb; }

int32_t dwba_semisep_lane_stage_group(int32_t L_max, int32_t lb) {
    lb = stage_lb(lb);
    if ((lb != 4 && lb != 8) || L_max < 0 || L_max > 63) return -1;
    const int nbp = lane_nbp(L_max, lb);
    return nbp < (lb == 8 ? 4 : 2) ? -1 : 64 / nbp;   // the copies of 32 streams at lb = 8 overflow vmcnt
}

int32_t dwba_il_semisep_lanes_stage_f64(int32_t n_jobs, const dwba_plain_job* d_jobs, const double* const* d_tabs,
                                        const int32_t* d_out_index, int32_t L_max, int32_t lb, double* d_out,
                                        void* stream) {
    DWBA_REQUIRE(initialised(), "dwba_il_semisep_lanes_stage_f64: call dwba_init first");
    lb = stage_lb(lb);
    DWBA_REQUIRE(n_jobs > 0 && d_jobs && d_tabs && d_out && dwba_semisep_lane_stage_group(L_max, lb) > 0,
                 "dwba_il_semisep_lanes_stage_f64: bad arguments (n_jobs=%d, L_max=%d, lb=%d; needs lb 0, 4 or 8 "
                 "and at least 2 multipole blocks)", n_jobs, L_max, lb);
    const int nbp = lane_nbp(L_max, lb);
    const long blocks = (2L * n_jobs * nbp + 127) / 128;
    DWBA_REQUIRE(blocks <= INT32_MAX, "dwba_il_semisep_lanes_stage_f64: too many jobs (%d)", n_jobs);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    switch (lb * 100 + nbp) {
#define C(B, P, NB, O) case B * 100 + P: \
        hipLaunchKernelGGL((semisep_lane_stage_kernel<B, P, NB, O>), dim3((unsigned)blocks), dim3(256), 0, s, d_jobs, \
                           d_tabs, d_out_index, n_jobs, L_max, d_out); break;
        // lb 4: 128-136 VGPRs, 4 wavefronts per SIMD, one chunk in flight; lb 8: 214 VGPRs, 2, two
        C(4, 2, 2, 4) C(4, 4, 2, 4) C(4, 8, 2, 4) C(4, 16, 2, 4) C(8, 4, 3, 2) C(8, 8, 3, 2)
#undef C
        default: return set_error(DWBA_ERR_ARG, "dwba_il_semisep_lanes_stage_f64: L_max %d, lb %d", L_max, lb);
    }
    DWBA_LAUNCH_CHECK("semisep_lane_stage_kernel");
    return DWBA_OK;
}

size_t dwba_il_backward_semisep_f64_workspace_bytes(int32_t n, int32_t n_pairs, int32_t L_max) {
    if (n < 2 || n_pairs < 1 || L_max < 0 || L_max > 63) return 0;
    return backward_layout(n, n_pairs, L_max).total;
}

int32_t dwba_il_backward_semisep_f64(const double* d_r, const double* d_w, int32_t n, int32_t n_pairs, int32_t ld,
                                     const double* d_rho_a, const double* d_rho_b, const double* d_g, int32_t L_max,
                                     double* d_grad_a, double* d_grad_b, void* d_workspace, size_t workspace_bytes,
                                     void* stream) {
    DWBA_REQUIRE(initialised(), "dwba_il_backward_semisep_f64: call dwba_init first");
    DWBA_REQUIRE(n >= 2 && n_pairs >= 1 && ld >= n && L_max >= 0 && L_max <= 63 && n_pairs <= 32767,
                 "dwba_il_backward_semisep_f64: bad shape (n=%d, pairs=%d, ld=%d, L_max=%d)", n, n_pairs, ld, L_max);
    DWBA_REQUIRE(d_r && d_w && d_rho_a && d_rho_b && d_g && d_workspace,
                 "dwba_il_backward_semisep_f64: null pointer");
    const BackwardLayout b = backward_layout(n, n_pairs, L_max);
    if (workspace_bytes < b.total)
        return set_error(DWBA_ERR_WORKSPACE, "dwba_il_backward_semisep_f64: workspace %zu < %zu bytes",
                         workspace_bytes, b.total);
    if (!d_grad_a && !d_grad_b) return DWBA_OK;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    char* ws = reinterpret_cast<char*>(d_workspace);
    auto* desc = reinterpret_cast<dwba_dgemv_energy*>(ws + b.desc_off);
    auto* xw = reinterpret_cast<double*>(ws + b.xw_off);
    auto* t = reinterpret_cast<double*>(ws + b.t_off);
    const int bx = std::min((n + GT - 1) / GT, 64);
    hipLaunchKernelGGL(backward_setup_kernel, dim3(bx, 2 * n_pairs), dim3(GT), 0, s, d_r, d_w, n, n_pairs, ld, d_rho_a,
                       d_rho_b, L_max, b.ld_t, xw, t, desc);
    DWBA_LAUNCH_CHECK("backward_setup_kernel");
    hipLaunchKernelGGL(semisep_gemv_scan_kernel, dim3(2 * n_pairs, L_max + 1), dim3(GT), 0, s, desc);
    DWBA_LAUNCH_CHECK("semisep_gemv_scan_kernel");
    const int cx = std::min((ld + GT - 1) / GT, 64);
    hipLaunchKernelGGL(backward_combine_kernel, dim3(cx, 2 * n_pairs), dim3(GT), 0, s, d_w, n, n_pairs, ld, d_g,
                       L_max, b.ld_t, t, d_grad_a, d_grad_b);
    DWBA_LAUNCH_CHECK("backward_combine_kernel");
    return DWBA_OK;
}

int32_t dwba_semisep_gemv_f64(int32_t n_energies, const dwba_dgemv_energy* d_energies, int32_t L_max,
                              void* stream) {
    DWBA_REQUIRE(initialised(), "dwba_semisep_gemv_f64: call dwba_init first");
    DWBA_REQUIRE(n_energies > 0 && d_energies && L_max >= 0 && L_max <= 63, "dwba_semisep_gemv_f64: bad arguments");
    hipLaunchKernelGGL(semisep_gemv_scan_kernel, dim3(n_energies, L_max + 1), dim3(GT), 0,
                       reinterpret_cast<hipStream_t>(stream), d_energies);
    DWBA_LAUNCH_CHECK("semisep_gemv_scan_kernel");
    return DWBA_OK;
}

}  // extern "C"
