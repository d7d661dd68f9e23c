// FP64 preparation of the alpha / beta vectors of one channel pair so that
// every quadrature branch of radial_ME_all_L becomes I_L = alpha^T K_L beta
// (dwba_matrix_elements.py:1030-1042, :1136-1245, :1281-1289 and
// oscillatory_kernel_integral_2d, oscillatory_integrals.py:2555-2777).
//
//  mode PLAIN (no oscillatory quadrature, or k_i + k_f <= k_threshold):
//     alpha_dir = (w_lim chi_f) chi_i, beta_dir = (w_lim u_f) u_i,
//     alpha_ex  = (w_lim u_f) chi_i,   beta_ex  = (w_lim chi_f) u_i,
//     w_lim = w on [0, N1), 0 beyond.
//  mode FILON: the outer Clenshaw-Curtis sum over phase intervals of
//     interp(rho1)(x_c) interp(K beta)(x_c) W_c is linear in K beta, hence
//     alpha_j = sum_c W_c interp(rho1)(x_c) lambda_c(j), lambda_c = the
//     interpolation weights of node c on r[:N1] (np.interp in log r on
//     log-spaced grids, :279-302).  beta_dir = (u_f u_i) w on the full grid;
//     beta_ex scatters W_c interp(chi_f u_i)(x_c) onto the two columns that
//     bracket x_c with the linear-in-r kernel interpolation weights of
//     :2729-2742.
//  mode STANDARD (Filon branch without usable phase intervals, :2596-2613):
//     alpha_* = rho1 w on [0, N1), beta_* = rho2 w on [0, N).
//
// The CC scatter is written as a gather -- grid point i visits the nodes of
// the phase intervals that overlap [r_{i-1}, r_{i+1}] in a fixed order -- so
// the result is deterministic without atomics.
//
// The same block also reduces the per-pair scalars: the L = 0 correction
// sums (:1249-1253, :1291-1296) and the multipole moments
// M_L = sum_j w_j r_j^L u_f u_i (:1167).
#include "common.hpp"
#include "prep.hpp"

namespace dwba {
namespace {

constexpr int PREP_THREADS = 256;

__device__ inline double clean(double v) { return isfinite(v) ? v : 0.0; }

// Phase node m of a job (uniform: numpy linspace; else the stored array).
__device__ inline double phase_node(const dwba_prep_job& J, int m) {
    if (J.phase_nodes) return J.phase_nodes[m];
    if (m == J.n_pnodes - 1) return J.p_end;
    return __dadd_rn(__dmul_rn((double)m, J.p_step), J.p_start);
}

// Interval index containing x (clamped to [0, n_int-1]).
__device__ inline int interval_of(const dwba_prep_job& J, double x, int n_int) {
    int m;
    if (!J.phase_nodes) {
        const double q = (x - J.p_start) / J.p_step;
        m = q <= 0.0 ? 0 : (q >= (double)n_int ? n_int - 1 : (int)q);
    } else {
        int lo = 0, hi = J.n_pnodes - 1;  // largest m with p[m] <= x
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (J.phase_nodes[mid] <= x) lo = mid; else hi = mid - 1;
        }
        m = lo;
    }
    return m < 0 ? 0 : (m > n_int - 1 ? n_int - 1 : m);
}

// np.interp cell on a table xp[0..n-1] (ascending): largest j <= n-2 with
// xp[j] <= X; X is already clipped to [xp[0], xp[n-1]].
__device__ inline int cell_of(const double* xp, int n, double X) {
    int lo = 0, hi = n - 2;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (xp[mid] <= X) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Value of np.interp(X, xp, fp) at cell j (slope form of numpy's kernel).
__device__ inline double interp_at(const double* xp, int n, double X, int j, double f0, double f1,
                                   double f_last) {
    if (X >= xp[n - 1]) return f_last;
    if (X == xp[j]) return f0;
    const double slope = (f1 - f0) / (xp[j + 1] - xp[j]);
    return slope * (X - xp[j]) + f0;
}

// Largest j in [0, n-2] with xp[j] <= X, X in [xp[0], xp[n-1]] (np.interp's
// cell).  The start is guessed from a uniform table and corrected by a short
// walk; a long walk (strongly non-uniform table) falls back to bisection.
__device__ inline int cell_fast(const double* xp, int n, double X) {
    if (n <= 2) return 0;
    const double x0 = xp[0], x1 = xp[n - 1];
    double f = (X - x0) / (x1 - x0) * (double)(n - 1);
    int j = f <= 0.0 ? 0 : (f >= (double)(n - 2) ? n - 2 : (int)f);
    for (int s = 0; s < 16; ++s) {
        if (xp[j] > X && j > 0) { --j; continue; }
        if (j < n - 2 && xp[j + 1] <= X) { ++j; continue; }
        return j;
    }
    return cell_of(xp, n, X);
}

// numpy searchsorted(r, x, 'left') clipped to [1, n-1], guessed in log r.
__device__ inline int right_col(const double* r, const double* lr, int n, double x) {
    int lo;   // number of r[j] < x
    if (!(r[0] < x)) lo = 0;
    else if (r[n - 1] < x) lo = n;
    else {
        // largest j with r[j] < x, in [0, n-2]
        const double f = (log(x) - lr[0]) / (lr[n - 1] - lr[0]) * (double)(n - 1);
        int j = f <= 0.0 ? 0 : (f >= (double)(n - 2) ? n - 2 : (int)f);
        int s = 0;
        for (; s < 16; ++s) {
            if (!(r[j] < x) && j > 0) { --j; continue; }
            if (j < n - 2 && r[j + 1] < x) { ++j; continue; }
            break;
        }
        if (s == 16) {
            int a = 0, b = n;
            while (a < b) {
                const int mid = (a + b) >> 1;
                if (r[mid] < x) a = mid + 1; else b = mid;
            }
            j = a - 1;
        }
        lo = j + 1;
    }
    return lo < 1 ? 1 : (lo > n - 1 ? n - 1 : lo);
}

struct CCTable {
    double x[DWBA_MAX_CC];
    double w[DWBA_MAX_CC];
    int n;
};

constexpr int CH_NODES = 512;    // CC nodes per LDS round

// FILON mode.  Every CC node c (interval m, node k, nodes ascending in x) is
// evaluated once per pair: its alpha cell ja (np.interp on r[:N1]) with the
// weights W f1(x) (1-t), W f1(x) t for direct and exchange, and its beta
// columns il = ir - 1 (searchsorted, linear-in-r kernel interpolation) with
// W f2(x) wl, W f2(x) wr.  Cells are non-decreasing along the node list, so
// a grid point finds the nodes that touch it by bisection in LDS and adds
// them in node order: deterministic, no atomics.
__device__ void filon_alpha_beta(const dwba_prep_job& J, const CCTable& cc) {
    __shared__ int s_ja[CH_NODES], s_jb[CH_NODES];
    __shared__ double s_ad0[CH_NODES], s_ad1[CH_NODES], s_ae0[CH_NODES], s_ae1[CH_NODES];
    __shared__ double s_b0[CH_NODES], s_b1[CH_NODES];
    const int N = J.N, N1 = J.N1;
    const int n_int = J.n_pnodes - 1;
    const int ncc = cc.n;
    const double* r = J.r;
    const double* lr = J.logr;                    // log(max(r, 1e-300))
    const double* xp1 = J.log1 ? lr : r;          // interpolation table on r[:N1]
    const double* xp2 = J.log2 ? lr : r;          // interpolation table on r[:N]
    const double r0 = r[0], r1e = r[N1 - 1], r2e = r[N - 1];
    const double rlog0 = r0 > 1e-300 ? r0 : 1e-300;
    const double rho_d_last = clean(J.chi_f[N1 - 1] * J.chi_i[N1 - 1]);
    const double rho_e_last = clean(J.u_f[N1 - 1] * J.chi_i[N1 - 1]);
    const double rho2_last = clean(J.chi_f[N - 1] * J.u_i[N - 1]);
    // Without DWBA_PREP_NO_ZERO_FILL the three vectors are cleared and every round
    // adds its contributions.  With it, only the points a round touches are
    // written: set on first touch, added where two consecutive rounds meet;
    // every other point keeps the zero it was allocated with (the same pair
    // touches the same points on every call).
    const bool fill = !(J.mode & DWBA_PREP_NO_ZERO_FILL);
    if (fill) {
        for (int i = threadIdx.x; i < N; i += blockDim.x) {
            J.alpha_dir[i] = 0.0;
            J.alpha_ex[i] = 0.0;
            J.beta_ex[i] = 0.0;
        }
    }
    int prev_ja_hi = -1, prev_jb_hi = -1;       // last points of the previous round's ranges
    const int per = CH_NODES / ncc;               // intervals per round
    for (int m0 = 0; m0 < n_int; m0 += per) {
        const int m_end = min(n_int, m0 + per);
        const int nq = (m_end - m0) * ncc;
        __syncthreads();                          // previous round's gathers are done
        for (int q = threadIdx.x; q < nq; q += blockDim.x) {
            const int m = m0 + q / ncc, k = q - (q / ncc) * ncc;
            const double a = phase_node(J, m), b = phase_node(J, m + 1);
            const double hw = 0.5 * (b - a);
            // clamped into [a, b] so that x never decreases along the node list
            const double x = fmin(fmax(hw * (cc.x[k] + 1.0) + a, a), b);
            const double W = (b - a > 1e-12) ? cc.w[k] * hw : 0.0;
            // alpha: np.interp of rho1 on r[:N1]
            const double xc = fmin(fmax(x, r0), r1e);
            const double X = J.log1 ? log(fmax(xc, rlog0)) : xc;
            const int j = cell_fast(xp1, N1, X);
            double t;
            if (X >= xp1[N1 - 1]) t = 1.0;
            else if (X == xp1[j]) t = 0.0;
            else t = (X - xp1[j]) / (xp1[j + 1] - xp1[j]);
            const double fd0 = clean(J.chi_f[j] * J.chi_i[j]), fd1 = clean(J.chi_f[j + 1] * J.chi_i[j + 1]);
            const double fe0 = clean(J.u_f[j] * J.chi_i[j]), fe1 = clean(J.u_f[j + 1] * J.chi_i[j + 1]);
            const double vd = W * interp_at(xp1, N1, X, j, fd0, fd1, rho_d_last);
            const double ve = W * interp_at(xp1, N1, X, j, fe0, fe1, rho_e_last);
            s_ja[q] = j;
            s_ad0[q] = vd * (1.0 - t);
            s_ad1[q] = vd * t;
            s_ae0[q] = ve * (1.0 - t);
            s_ae1[q] = ve * t;
            // beta_ex: kernel interpolation linear in r between columns il, ir
            const int ir = right_col(r, lr, N, x);
            const int il = ir - 1;
            const double wr = (x - r[il]) / (r[ir] - r[il] + 1e-30);
            const double wl = 1.0 - wr;
            const double xc2 = fmin(fmax(x, r0), r2e);
            const double X2 = J.log2 ? log(fmax(xc2, rlog0)) : xc2;
            const int j2 = cell_fast(xp2, N, X2);
            const double f0 = clean(J.chi_f[j2] * J.u_i[j2]), f1 = clean(J.chi_f[j2 + 1] * J.u_i[j2 + 1]);
            const double v = W * interp_at(xp2, N, X2, j2, f0, f1, rho2_last);
            s_jb[q] = il;
            s_b0[q] = v * wl;
            s_b1[q] = v * wr;
        }
        __syncthreads();
        // gather onto the grid points this round touches
        const int ja_lo = s_ja[0], ja_hi = s_ja[nq - 1] + 1;
        const int jb_lo = s_jb[0], jb_hi = s_jb[nq - 1] + 1;
        for (int i = ja_lo + (int)threadIdx.x; i <= ja_hi; i += blockDim.x) {
            // nodes with ja == i-1 contribute their t-part, nodes with ja == i their (1-t)-part
            int lo = 0, hi = nq;                  // first q with ja >= i-1
            while (lo < hi) { const int mid = (lo + hi) >> 1; if (s_ja[mid] < i - 1) lo = mid + 1; else hi = mid; }
            double sd = 0.0, se = 0.0;
            int q = lo;
            for (; q < nq && s_ja[q] == i - 1; ++q) { sd += s_ad1[q]; se += s_ae1[q]; }
            for (; q < nq && s_ja[q] == i; ++q) { sd += s_ad0[q]; se += s_ae0[q]; }
            if (i < N1) {
                if (fill) {
                    if (sd != 0.0 || se != 0.0) {
                        J.alpha_dir[i] += sd;
                        J.alpha_ex[i] += se;
                    }
                } else if (i <= prev_ja_hi) {
                    J.alpha_dir[i] += sd;
                    J.alpha_ex[i] += se;
                } else {
                    J.alpha_dir[i] = sd;
                    J.alpha_ex[i] = se;
                }
            }
        }
        for (int i = jb_lo + (int)threadIdx.x; i <= jb_hi; i += blockDim.x) {
            int lo = 0, hi = nq;
            while (lo < hi) { const int mid = (lo + hi) >> 1; if (s_jb[mid] < i - 1) lo = mid + 1; else hi = mid; }
            double sb = 0.0;
            int q = lo;
            for (; q < nq && s_jb[q] == i - 1; ++q) sb += s_b1[q];
            for (; q < nq && s_jb[q] == i; ++q) sb += s_b0[q];
            if (i < N) {
                if (fill) {
                    if (sb != 0.0) J.beta_ex[i] += sb;
                } else if (i <= prev_jb_hi) {
                    J.beta_ex[i] += sb;
                } else {
                    J.beta_ex[i] = sb;
                }
            }
        }
        prev_ja_hi = ja_hi;
        prev_jb_hi = jb_hi;
    }
}

__global__ __launch_bounds__(PREP_THREADS) void prep_f64_kernel(const dwba_prep_job* __restrict__ jobs,
                                                                CCTable cc) {
    const dwba_prep_job J = jobs[blockIdx.x];
    const int N = J.N, N1 = J.N1;
    __shared__ double s_red[PREP_THREADS / 64][4];

    // ---------------- one streaming pass over the grid ----------------
    // L = 0 correction factors over the whole grid with w_lim = 0 beyond N1,
    // as the reference's dot products do (a non-finite density beyond N1 still
    // propagates), and every output that is a pointwise product of the inputs.
    const int mode = J.mode & DWBA_PREP_MODE_MASK;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
        const double ui = J.u_i[i], uf = J.u_f[i], ci = J.chi_i[i], cf = J.chi_f[i];
        const double wf = J.w[i];
        const double w = i < N1 ? wf : 0.0;
        const double vd = J.V_diff ? J.V_diff[i] : 0.0;
        acc[0] += (w * uf) * ui;              // rho2_dir_w
        acc[1] += ((w * cf) * ci) * vd;       // rho1_dir_w . V_diff
        acc[2] += (w * cf) * ui;              // rho2_ex_w
        acc[3] += ((w * uf) * ci) * vd;       // rho1_ex_w . V_diff
        if (mode == DWBA_PREP_PLAIN) {
            J.alpha_dir[i] = (w * cf) * ci;
            J.beta_dir[i] = (w * uf) * ui;
            J.alpha_ex[i] = (w * uf) * ci;
            J.beta_ex[i] = (w * cf) * ui;
        } else if (mode == DWBA_PREP_ZERO) {
            J.alpha_dir[i] = 0.0; J.beta_dir[i] = 0.0; J.alpha_ex[i] = 0.0; J.beta_ex[i] = 0.0;
        } else {
            // beta_dir = (u_f u_i) w on the full grid for both STANDARD and FILON
            J.beta_dir[i] = clean(uf * ui) * wf;
            if (mode == DWBA_PREP_STANDARD) {
                J.alpha_dir[i] = clean(cf * ci) * w;
                J.alpha_ex[i] = clean(uf * ci) * w;
                J.beta_ex[i] = clean(cf * ui) * wf;
            }
        }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const double sum = wave_sum(acc[q]);
        if (lane == 0) s_red[wave][q] = sum;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        double sum = 0.0;
        for (int wv = 0; wv < PREP_THREADS / 64; ++wv) sum += s_red[wv][threadIdx.x];
        J.scalars[threadIdx.x] = sum;
    }
    if (mode != DWBA_PREP_FILON) return;

    // FILON
    filon_alpha_beta(J, cc);
}

// Multipole moments: block (set, L) reduces sum_j ((w_j r_j^L) u_f,j) u_i,j.
__global__ __launch_bounds__(PREP_THREADS) void moments_f64_kernel(const double* const* __restrict__ R,
                                                                   const double* const* __restrict__ Wt,
                                                                   const double* const* __restrict__ UI,
                                                                   const double* const* __restrict__ UF,
                                                                   const int* __restrict__ NN, int L_max,
                                                                   double* __restrict__ out) {
    __shared__ double s_red[PREP_THREADS / 64];
    const int set = blockIdx.x, L = blockIdx.y;
    const double *r = R[set], *w = Wt[set], *ui = UI[set], *uf = UF[set];
    const int n = NN[set];
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const double rl = L == 0 ? 1.0 : (L == 1 ? r[i] : (L == 2 ? r[i] * r[i] : pow(r[i], (double)L)));
        acc += ((w[i] * rl) * uf[i]) * ui[i];
    }
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double sum = 0.0;
        for (int k = 0; k < PREP_THREADS / 64; ++k) sum += s_red[k];
        out[(size_t)set * (L_max + 1) + L] = sum;
    }
}

// Coulomb-phase node stepping of generate_phase_nodes (oscillatory_integrals.py:1685-1713):
// r_{n+1} = r_n + clip(dphi / |k + eta / max(r_n, 1e-10)|, 1e-6, (b - a)/10), capped at b,
// at most 10001 nodes.  One thread per job; node count written to the job.
__global__ void coulomb_nodes_kernel(dwba_prep_job* __restrict__ jobs, int n_jobs, const double* __restrict__ par) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_jobs) return;
    dwba_prep_job& J = jobs[k];
    if ((J.mode & DWBA_PREP_MODE_MASK) != DWBA_PREP_FILON || J.phase_nodes == nullptr) return;
    const double a = par[4 * k + 0], b = par[4 * k + 1], kt = par[4 * k + 2], eta = par[4 * k + 3];
    const double dphi = J.p_step;  // phase increment for Coulomb jobs
    double* nodes = const_cast<double*>(J.phase_nodes);
    int n = 0;
    nodes[n++] = a;
    double r = a;
    const double cap = (b - a) / 10;
    while (r < b) {
        const double pp = kt + eta / fmax(r, 1e-10);
        double dr = dphi / fabs(pp);
        dr = fmin(dr, cap);
        dr = fmax(dr, 1e-6);
        double rn = r + dr;
        if (rn >= b) rn = b;
        nodes[n++] = rn;
        r = rn;
        if (n > 10000) break;
    }
    J.n_pnodes = n;
}

}  // namespace

int launch_prep(int n_jobs, dwba_prep_job* d_jobs, const double* d_coulomb_par, int any_coulomb,
                const double* cc_x, const double* cc_w, int n_cc, hipStream_t s) {
    CCTable cc{};
    if (n_cc < 2 || n_cc > DWBA_MAX_CC) return set_error(DWBA_ERR_ARG, "prep: CC node count %d outside [2, %d]", n_cc, DWBA_MAX_CC);
    // nodes ascending in x (the FILON gather relies on it; the sum order is free)
    int ord[DWBA_MAX_CC];
    for (int k = 0; k < n_cc; ++k) ord[k] = k;
    std::sort(ord, ord + n_cc, [&](int p, int q) { return cc_x[p] < cc_x[q]; });
    for (int k = 0; k < n_cc; ++k) { cc.x[k] = cc_x[ord[k]]; cc.w[k] = cc_w[ord[k]]; }
    cc.n = n_cc;
    if (any_coulomb) {
        hipLaunchKernelGGL(coulomb_nodes_kernel, dim3((n_jobs + 63) / 64), dim3(64), 0, s, d_jobs, n_jobs, d_coulomb_par);
        DWBA_LAUNCH_CHECK("coulomb_nodes_kernel");
    }
    hipLaunchKernelGGL(prep_f64_kernel, dim3(n_jobs), dim3(PREP_THREADS), 0, s, d_jobs, cc);
    DWBA_LAUNCH_CHECK("prep_f64_kernel");
    return DWBA_OK;
}

}  // namespace dwba

extern "C" int32_t dwba_moments_f64(int32_t n_sets, const double* const* d_r, const double* const* d_w,
                                    const double* const* d_u_i, const double* const* d_u_f, const int32_t* d_n,
                                    int32_t L_max, double* d_out, void* stream) {
    DWBA_REQUIRE(dwba::initialised(), "dwba_moments_f64: call dwba_init first");
    DWBA_REQUIRE(n_sets > 0 && L_max >= 0 && L_max < 65535 && d_r && d_w && d_u_i && d_u_f && d_n && d_out,
                 "dwba_moments_f64: bad arguments");
    hipLaunchKernelGGL(dwba::moments_f64_kernel, dim3(n_sets, L_max + 1), dim3(dwba::PREP_THREADS), 0,
                       reinterpret_cast<hipStream_t>(stream), d_r, d_w, d_u_i, d_u_f, d_n, L_max, d_out);
    DWBA_LAUNCH_CHECK("moments_f64_kernel");
    return DWBA_OK;
}

extern "C" int32_t dwba_fullsplit_prep_f64(int32_t n_jobs, dwba_prep_job* d_jobs, const double* d_coulomb_par,
                                           int32_t any_coulomb, const double* h_cc_x, const double* h_cc_w,
                                           int32_t n_cc, void* stream) {
    DWBA_REQUIRE(dwba::initialised(), "dwba_fullsplit_prep_f64: call dwba_init first");
    DWBA_REQUIRE(n_jobs > 0 && d_jobs && h_cc_x && h_cc_w, "dwba_fullsplit_prep_f64: bad arguments");
    DWBA_REQUIRE(!any_coulomb || d_coulomb_par, "dwba_fullsplit_prep_f64: Coulomb jobs need d_coulomb_par");
    return dwba::launch_prep(n_jobs, d_jobs, d_coulomb_par, any_coulomb, h_cc_x, h_cc_w, n_cc,
                             reinterpret_cast<hipStream_t>(stream));
}
