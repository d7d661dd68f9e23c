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

struct CCTable {
    double x[DWBA_MAX_CC];
    double w[DWBA_MAX_CC];
    int n;
};

// Ends [a, b] of phase interval m < n_pnodes - 1: the stored nodes (Coulomb
// stepping) or numpy linspace(p_start, p_end, n_pnodes) (uniform).
template <bool COUL>
__device__ inline void interval_ends(const dwba_prep_job& J, int m, double& a, double& b) {
    if (COUL) {
        a = J.phase_nodes[m];
        b = J.phase_nodes[m + 1];
    } else {
        a = __dadd_rn(__dmul_rn((double)m, J.p_step), J.p_start);
        const double bn = __dadd_rn(__dmul_rn((double)(m + 1), J.p_step), J.p_start);
        b = m + 1 == J.n_pnodes - 1 ? J.p_end : bn;
    }
}

// Flattened index g = m ncc + k of the first CC node with x >= v (n_int ncc
// when there is none).  Node k of interval m sits at x = clamp(hw (c_k + 1)
// + a, a, b), hw = (b - a)/2, c_k ascending, so x never decreases along g.
// Starts one interval before the one holding v, so a rounded interval index
// never skips a node.
template <bool COUL>
__device__ inline int first_node_ge(const dwba_prep_job& J, const double* ccx, int ncc, int n_int, double v) {
    int m = interval_of(J, v, n_int) - 1;
    for (m = m < 0 ? 0 : m; m < n_int; ++m) {
        double a, b;
        interval_ends<COUL>(J, m, a, b);
        if (b < v) continue;                      // every node of m is <= b < v
        const double hw = 0.5 * (b - a);
        for (int k = 0; k < ncc; ++k)
            if (fmin(fmax(hw * (ccx[k] + 1.0) + a, a), b) >= v) return m * ncc + k;
    }
    return n_int * ncc;
}

// log(1 + u) for 0 <= u < 0.01: series to u^9 (truncation below 1e-19 u).
__device__ inline double log1p_small(double u) {
    double p = 1.0 / 9.0;
    p = fma(p, u, -1.0 / 8.0);
    p = fma(p, u, 1.0 / 7.0);
    p = fma(p, u, -1.0 / 6.0);
    p = fma(p, u, 1.0 / 5.0);
    p = fma(p, u, -1.0 / 4.0);
    p = fma(p, u, 1.0 / 3.0);
    p = fma(p, u, -1.0 / 2.0);
    p = fma(p, u, 1.0);
    return p * u;
}

constexpr int NODE_BATCH = 4;    // CC nodes evaluated side by side per cell

// FILON mode as one pass over the grid cells [r_j, r_{j+1}) of r[:N1].
// A CC node c at x in cell j (interval m, node k, weight W = w_k (b - a)/2)
// adds
//   alpha: W interp(rho1)(x) (1 - t), W interp(rho1)(x) t    to points j, j+1,
//          interp(rho1)(x) = (1 - t) rho1_j + t rho1_{j+1}, t the np.interp
//          weight in log r or r;
//   beta_ex: W interp(rho2)(x) wl, W interp(rho2)(x) wr      to columns j, j+1,
//          wl, wr the linear-in-r kernel weights.
// Summed over the nodes of the cell this is a 2x2 operator on the densities
// at j, j+1 whose coefficients (sums of W t^a (1-t)^b) need no density, so a
// thread owns one cell, finds the index range of its nodes from their
// positions (the upper end is the next cell's lower end, passed through LDS)
// and evaluates them NODE_BATCH at a time as independent chains; point i
// takes the cell-i part plus the cell-(i-1) part from its neighbour.  Nodes
// lie in [r_0, r_{N1-1}]; the last cell is closed on the right.  Every node
// lands in exactly one cell; because np.interp and the kernel weights are
// continuous, a node on a cell edge gives the same contribution (to
// rounding) from either side.  In log r the weight is
// t = log(x / r_j) / (log r_{j+1} - log r_j); the numerator is the series
// log1p((x - r_j) / r_j) in cells narrower than 1% of r_j (it differs from
// the rounded difference log x - log r_j by less than that difference's own
// rounding error).
// The 2x2 interpolation sums of cell j's CC nodes g_lo .. g_hi - 1 (see
// filon_alpha_beta below), evaluated NODE_BATCH nodes at a time.
template <bool COUL>
__device__ __forceinline__ void cell_sums(const dwba_prep_job& J, const double* s_ccx, const double* s_ccw, int ncc,
                                          double inv_ncc, bool lg1, bool lg2, int j, int g_lo, int g_hi,
                                          double& c00, double& c01, double& c11, double& b00, double& b01,
                                          double& b10, double& b11) {
    const double* r = J.r;
    const double* lr = J.logr;
    const bool need_log = lg1 || lg2;
    const double rl = r[j], rr = r[j + 1];
    const double ll = lr[j], lrr = lr[j + 1];
    const double inv_log = lrr > ll ? 1.0 / (lrr - ll) : 0.0;
    const double inv_lin = rr > rl ? 1.0 / (rr - rl) : 0.0;
    const double invr = 1.0 / (rr - rl + 1e-30);
    const bool ser = rl > 0.0 && (rr - rl) < 0.01 * rl;
    const double inv_rl = ser ? 1.0 / rl : 0.0;
    c00 = 0.0; c01 = 0.0; c11 = 0.0; b00 = 0.0; b01 = 0.0; b10 = 0.0; b11 = 0.0;
    for (int g0 = g_lo; g0 < g_hi; g0 += NODE_BATCH) {
        double xs[NODE_BATCH], Ws[NODE_BATCH], dX[NODE_BATCH];
#pragma unroll
        for (int v = 0; v < NODE_BATCH; ++v) {
            const int g = min(g0 + v, g_hi - 1);  // a node past the range repeats with weight 0
            const int m = (int)(((double)g + 0.5) * inv_ncc);
            const int k = g - m * ncc;
            double a, b;
            interval_ends<COUL>(J, m, a, b);
            const double hw = 0.5 * (b - a);
            const double W = s_ccw[k] * hw;
            Ws[v] = (b - a > 1e-12 && g0 + v < g_hi) ? W : 0.0;
            xs[v] = fmin(fmax(hw * (s_ccx[k] + 1.0) + a, a), b);
            dX[v] = 0.0;
        }
        if (need_log) {
            if (ser) {
#pragma unroll
                for (int v = 0; v < NODE_BATCH; ++v) dX[v] = log1p_small((xs[v] - rl) * inv_rl);
            } else {
#pragma unroll
                for (int v = 0; v < NODE_BATCH; ++v) dX[v] = log(fmax(xs[v], 1e-300)) - ll;
            }
        }
#pragma unroll
        for (int v = 0; v < NODE_BATCH; ++v) {
            const double W = Ws[v];
            const double dx = xs[v] - rl;
            const double tl = dX[v] * inv_log, tn = dx * inv_lin;
            const double t1 = fmin(fmax(lg1 ? tl : tn, 0.0), 1.0);
            const double t2 = fmin(fmax(lg2 ? tl : tn, 0.0), 1.0);
            const double wr = dx * invr, wl = 1.0 - wr;
            const double u1 = 1.0 - t1, u2 = 1.0 - t2;
            const double Wu = W * u1, Wt = W * t1;
            c00 = fma(Wu, u1, c00);
            c01 = fma(Wu, t1, c01);
            c11 = fma(Wt, t1, c11);
            const double Wl = W * wl, Wr = W * wr;
            b00 = fma(Wl, u2, b00);
            b01 = fma(Wl, t2, b01);
            b10 = fma(Wr, u2, b10);
            b11 = fma(Wr, t2, b11);
        }
    }
}

template <bool COUL>
__device__ void filon_alpha_beta(const dwba_prep_job& J, const CCTable& cc) {
    __shared__ double s_s1[PREP_THREADS][3];
    __shared__ double s_ccx[DWBA_MAX_CC], s_ccw[DWBA_MAX_CC];
    __shared__ int s_glo[PREP_THREADS];
    const int N = J.N, N1 = J.N1;
    const int n_int = J.n_pnodes - 1;
    const int ncc = cc.n;
    const int g_end = n_int * ncc;
    const double inv_ncc = 1.0 / ncc;
    const double* r = J.r;
    const bool lg1 = J.log1 != 0, lg2 = J.log2 != 0;
    const bool fill = !(J.mode & DWBA_PREP_NO_ZERO_FILL);
    if (threadIdx.x < DWBA_MAX_CC) {              // per-lane node indices read the table from LDS
        s_ccx[threadIdx.x] = threadIdx.x < ncc ? cc.x[threadIdx.x] : 0.0;
        s_ccw[threadIdx.x] = threadIdx.x < ncc ? cc.w[threadIdx.x] : 0.0;
    }
    __syncthreads();
    double carry_d = 0.0, carry_e = 0.0, carry_b = 0.0;   // cell-(i-1) parts across chunks
    for (int base = 0; base < N1; base += PREP_THREADS) {
        const int j = base + (int)threadIdx.x;
        const bool cell = j < N1 - 1;
        const bool last = j == N1 - 2;
        const int g_lo = !cell ? g_end : (j == 0 ? 0 : first_node_ge<COUL>(J, s_ccx, ncc, n_int, r[j]));
        s_glo[threadIdx.x] = g_lo;
        __syncthreads();
        double s0d = 0.0, s1d = 0.0, s0e = 0.0, s1e = 0.0, s0b = 0.0, s1b = 0.0;
        if (cell) {
            const int g_hi = last ? g_end
                                  : (threadIdx.x + 1 < PREP_THREADS ? s_glo[threadIdx.x + 1]
                                                                    : first_node_ge<COUL>(J, s_ccx, ncc, n_int, r[j + 1]));
            double c00, c01, c11, b00, b01, b10, b11;
            cell_sums<COUL>(J, s_ccx, s_ccw, ncc, inv_ncc, lg1, lg2, j, g_lo, g_hi, c00, c01, c11, b00, b01, b10, b11);
            const double ci0 = J.chi_i[j], ci1 = J.chi_i[j + 1];
            const double d0 = clean(J.chi_f[j] * ci0), d1 = clean(J.chi_f[j + 1] * ci1);
            const double e0 = clean(J.u_f[j] * ci0), e1 = clean(J.u_f[j + 1] * ci1);
            const double f0 = clean(J.chi_f[j] * J.u_i[j]), f1 = clean(J.chi_f[j + 1] * J.u_i[j + 1]);
            s0d = c00 * d0 + c01 * d1;
            s1d = c01 * d0 + c11 * d1;
            s0e = c00 * e0 + c01 * e1;
            s1e = c01 * e0 + c11 * e1;
            s0b = b00 * f0 + b01 * f1;
            s1b = b10 * f0 + b11 * f1;
        }
        s_s1[threadIdx.x][0] = s1d;
        s_s1[threadIdx.x][1] = s1e;
        s_s1[threadIdx.x][2] = s1b;
        __syncthreads();
        if (j < N1) {
            const bool first = threadIdx.x == 0;
            J.alpha_dir[j] = s0d + (first ? carry_d : s_s1[threadIdx.x - 1][0]);
            J.alpha_ex[j] = s0e + (first ? carry_e : s_s1[threadIdx.x - 1][1]);
            J.beta_ex[j] = s0b + (first ? carry_b : s_s1[threadIdx.x - 1][2]);
        }
        carry_d = s_s1[PREP_THREADS - 1][0];
        carry_e = s_s1[PREP_THREADS - 1][1];
        carry_b = s_s1[PREP_THREADS - 1][2];
        __syncthreads();                          // s_s1 is rewritten by the next chunk
    }
    // Without DWBA_PREP_NO_ZERO_FILL the points beyond N1 (no node reaches
    // them) are cleared; with it they keep the zeros they were allocated with.
    if (fill) {
        for (int i = N1 + (int)threadIdx.x; i < N; i += PREP_THREADS) {
            J.alpha_dir[i] = 0.0;
            J.alpha_ex[i] = 0.0;
            J.beta_ex[i] = 0.0;
        }
    }
}

__global__ __launch_bounds__(PREP_THREADS) void prep_f64_kernel(const dwba_prep_job* __restrict__ jobs,
                                                                CCTable cc) {
    const dwba_prep_job J = jobs[blockIdx.x];
    if (J.mode & DWBA_PREP_FILON_GROUPED) return;   // dwba_filon_cells_f64 + dwba_filon_fused_f64 do it all
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
            if (J.beta_dir) J.beta_dir[i] = clean(uf * ui) * wf;   // NULL: a shared beta written by another pair
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
    if (J.phase_nodes) filon_alpha_beta<true>(J, cc);
    else filon_alpha_beta<false>(J, cc);
}

// Cell sums of a group of FILON jobs sharing their geometry: one workgroup
// per group, the cell pass of filon_alpha_beta without densities; the 7
// sums of cell j go to coef[q * ld_coef + j].
template <bool COUL>
__device__ void filon_cells(const dwba_filon_group& G, const dwba_prep_job& J, const CCTable& cc) {
    __shared__ double s_ccx[DWBA_MAX_CC], s_ccw[DWBA_MAX_CC];
    __shared__ int s_glo[PREP_THREADS];
    const int N1 = J.N1;
    const int n_int = J.n_pnodes - 1;
    const int ncc = cc.n;
    const int g_end = n_int * ncc;
    const double inv_ncc = 1.0 / ncc;
    const double* r = J.r;
    const bool lg1 = J.log1 != 0, lg2 = J.log2 != 0;
    if (threadIdx.x < DWBA_MAX_CC) {
        s_ccx[threadIdx.x] = threadIdx.x < ncc ? cc.x[threadIdx.x] : 0.0;
        s_ccw[threadIdx.x] = threadIdx.x < ncc ? cc.w[threadIdx.x] : 0.0;
    }
    __syncthreads();
    double* __restrict__ co = G.coef;
    const size_t ld = (size_t)G.ld_coef;
    for (int base = 0; base < N1; base += PREP_THREADS) {
        const int j = base + (int)threadIdx.x;
        const bool cell = j < N1 - 1;
        const bool last = j == N1 - 2;
        const int g_lo = !cell ? g_end : (j == 0 ? 0 : first_node_ge<COUL>(J, s_ccx, ncc, n_int, r[j]));
        s_glo[threadIdx.x] = g_lo;
        __syncthreads();
        if (cell) {
            const int g_hi = last ? g_end
                                  : (threadIdx.x + 1 < PREP_THREADS ? s_glo[threadIdx.x + 1]
                                                                    : first_node_ge<COUL>(J, s_ccx, ncc, n_int, r[j + 1]));
            double c00, c01, c11, b00, b01, b10, b11;
            cell_sums<COUL>(J, s_ccx, s_ccw, ncc, inv_ncc, lg1, lg2, j, g_lo, g_hi, c00, c01, c11, b00, b01, b10, b11);
            co[j] = c00;
            co[ld + j] = c01;
            co[2 * ld + j] = c11;
            co[3 * ld + j] = b00;
            co[4 * ld + j] = b01;
            co[5 * ld + j] = b10;
            co[6 * ld + j] = b11;
        }
        __syncthreads();                          // s_glo is rewritten by the next chunk
    }
}

__global__ __launch_bounds__(PREP_THREADS) void filon_cells_kernel(const dwba_filon_group* __restrict__ groups,
                                                                   const dwba_prep_job* __restrict__ jobs,
                                                                   CCTable cc) {
    const dwba_filon_group G = groups[blockIdx.x];
    const dwba_prep_job J = jobs[G.rep];
    if (J.phase_nodes) filon_cells<true>(G, J, cc);
    else filon_cells<false>(G, J, cc);
}

// Grouped Filon prep of one slice of <= FUSE_MAX members of a group (same
// grid, split index, phase layout and interpolation kinds), fused with the
// per-pair streaming pass: one workgroup per slice sweeps the grid [0, N)
// in chunks of 256 points, a thread per point.  The group's cell sums of
// cell j and (from the neighbouring lane) cell j - 1, and the grid data
// shared by the members (u_i, w, V_diff) are loaded once per point; each
// member's densities once per member and point, the values at j - 1 and
// j + 1 coming from the neighbouring lanes (lanes 0 and 63 load theirs).
// Per member it writes, for j < N1,
//   alpha_dir[j] = (c00 d_j + c01 d_j+1) + (c01' d_j-1 + c11' d_j),
//   alpha_ex[j]  the same with e,  beta_ex[j] = (b00 f_j + b01 f_j+1) + (b10' f_j-1 + b11' f_j)
// (primed: cell j - 1; d = chi_f chi_i, e = u_f chi_i, f = chi_f u_i, non-finite
// values cleared) -- the per-pair cell pass's sums -- and beta_dir = (u_f u_i) w
// on [0, N) where the job has a beta_dir, and reduces the four L = 0
// correction sums of the streaming pass over the whole grid with w_lim = 0
// beyond N1 (a non-finite density there still propagates, as in the
// reference's dot products).  Workgroups are dealt to the 8 XCDs
// round-robin; slice s runs on XCD s / ceil(n / 8), so a caller that lists
// its slices in eight consecutive runs (one energy's slices in one run,
// count 0 for padding) keeps each energy's shared arrays and cell tables in
// one XCD's L2.
constexpr int FUSE_MAX = 4;

__global__ __launch_bounds__(PREP_THREADS) void filon_fused_kernel(const dwba_filon_group* __restrict__ slices,
                                                                   const dwba_prep_job* __restrict__ jobs,
                                                                   const int* __restrict__ members) {
    __shared__ double s_red[PREP_THREADS / 64][FUSE_MAX * 4];
    const int nb = gridDim.x, b = blockIdx.x;
    const int per = nb / 8, rem = nb % 8, xcd = b % 8;
    const int sidx = (xcd < rem ? xcd * (per + 1) : rem * (per + 1) + (xcd - rem) * per) + b / 8;
    if (sidx >= nb) return;
    const dwba_filon_group G = slices[sidx];
    if (G.count <= 0) return;
    const dwba_prep_job& R = jobs[G.rep];
    const int N = R.N, N1 = R.N1;
    const int cnt = G.count < FUSE_MAX ? G.count : FUSE_MAX;
    // points at and beyond N1 matter only for a member that writes beta_dir
    // there or whose densities may be non-finite there (w_lim = 0 makes every
    // L = 0 term an exact zero otherwise)
    int nend = N1;
    for (int m = 0; m < cnt; ++m) {
        const dwba_prep_job& J = jobs[members[G.first + m]];
        if (J.beta_dir || !(J.mode & DWBA_PREP_FINITE_TAIL)) nend = N;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const size_t ld = (size_t)G.ld_coef;
    const double* __restrict__ co = G.coef;
    const double* __restrict__ ui_p = R.u_i;
    const double* __restrict__ w_p = R.w;
    const double* __restrict__ v_p = R.V_diff;
    const dwba_prep_job* Jm[FUSE_MAX];
#pragma unroll
    for (int m = 0; m < FUSE_MAX; ++m) Jm[m] = &jobs[members[G.first + (m < cnt ? m : 0)]];
    double acc[FUSE_MAX][4];
#pragma unroll
    for (int m = 0; m < FUSE_MAX; ++m)
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[m][q] = 0.0;
    for (int base = 0; base < nend; base += PREP_THREADS) {
        const int j = base + (int)threadIdx.x;
        const bool in = j < N;
        const bool here = j < N1 - 1;                   // cell j exists
        const bool prev = j >= 1 && j < N1;             // cell j - 1 exists
        // ---- every load of the chunk first (the stores below could alias them)
        double c00 = 0.0, c01 = 0.0, c11 = 0.0, b00 = 0.0, b01 = 0.0, b10 = 0.0, b11 = 0.0;
        if (here) {
            c00 = co[j]; c01 = co[ld + j]; c11 = co[2 * ld + j];
            b00 = co[3 * ld + j]; b01 = co[4 * ld + j]; b10 = co[5 * ld + j]; b11 = co[6 * ld + j];
        }
        double p01 = 0.0, p11 = 0.0, q10 = 0.0, q11 = 0.0;   // cell j - 1 (lane 0: loaded)
        if (lane == 0 && prev) { p01 = co[ld + j - 1]; p11 = co[2 * ld + j - 1]; q10 = co[5 * ld + j - 1]; q11 = co[6 * ld + j - 1]; }
        const double ui = in ? ui_p[j] : 0.0;
        const double wf = in ? w_p[j] : 0.0;
        const double vd = (in && v_p) ? v_p[j] : 0.0;
        // the neighbours a wave's end lanes need: lane 63 at j + 1, lane 0 at j - 1
        const int jh = lane == 63 ? j + 1 : j - 1;
        const bool halo = (lane == 63 && here) || (lane == 0 && prev);
        const double uih = halo ? ui_p[jh] : 0.0;
        double ci[FUSE_MAX], cf[FUSE_MAX], uf[FUSE_MAX], cih[FUSE_MAX], cfh[FUSE_MAX], ufh[FUSE_MAX];
#pragma unroll
        for (int m = 0; m < FUSE_MAX; ++m) {
            const bool live = in && m < cnt;
            ci[m] = live ? Jm[m]->chi_i[j] : 0.0;
            cf[m] = live ? Jm[m]->chi_f[j] : 0.0;
            uf[m] = live ? Jm[m]->u_f[j] : 0.0;
            const bool hl = halo && m < cnt;
            cih[m] = hl ? Jm[m]->chi_i[jh] : 0.0;
            cfh[m] = hl ? Jm[m]->chi_f[jh] : 0.0;
            ufh[m] = hl ? Jm[m]->u_f[jh] : 0.0;
        }
        {
            const double t01 = __shfl_up(c01, 1, 64), t11 = __shfl_up(c11, 1, 64);
            const double t10 = __shfl_up(b10, 1, 64), u11 = __shfl_up(b11, 1, 64);
            if (lane != 0) { p01 = t01; p11 = t11; q10 = t10; q11 = u11; }
        }
        const double w = j < N1 ? wf : 0.0;
        // ---- compute and store
#pragma unroll
        for (int m = 0; m < FUSE_MAX; ++m) {
            if (m >= cnt) break;
            const dwba_prep_job& J = *Jm[m];
            acc[m][0] += (w * uf[m]) * ui;            // rho2_dir_w
            acc[m][1] += ((w * cf[m]) * ci[m]) * vd;  // rho1_dir_w . V_diff
            acc[m][2] += (w * cf[m]) * ui;            // rho2_ex_w
            acc[m][3] += ((w * uf[m]) * ci[m]) * vd;  // rho1_ex_w . V_diff
            if (in && J.beta_dir) J.beta_dir[j] = clean(uf[m] * ui) * wf;
            const double d = clean(cf[m] * ci[m]), e = clean(uf[m] * ci[m]), f = clean(cf[m] * ui);
            double d1 = __shfl_down(d, 1, 64), e1 = __shfl_down(e, 1, 64), f1 = __shfl_down(f, 1, 64);
            double d0 = __shfl_up(d, 1, 64), e0 = __shfl_up(e, 1, 64), f0 = __shfl_up(f, 1, 64);
            const double dh = clean(cfh[m] * cih[m]), eh = clean(ufh[m] * cih[m]), fh = clean(cfh[m] * uih);
            if (lane == 63) { d1 = dh; e1 = eh; f1 = fh; }
            if (lane == 0) { d0 = dh; e0 = eh; f0 = fh; }
            if (j < N1) {
                double s0d = 0.0, s0e = 0.0, s0b = 0.0, s1d = 0.0, s1e = 0.0, s1b = 0.0;
                if (here) {
                    s0d = c00 * d + c01 * d1;
                    s0e = c00 * e + c01 * e1;
                    s0b = b00 * f + b01 * f1;
                }
                if (prev) {
                    s1d = p01 * d0 + p11 * d;
                    s1e = p01 * e0 + p11 * e;
                    s1b = q10 * f0 + q11 * f;
                }
                J.alpha_dir[j] = s0d + s1d;
                J.alpha_ex[j] = s0e + s1e;
                J.beta_ex[j] = s0b + s1b;
            }
        }
    }
#pragma unroll
    for (int m = 0; m < FUSE_MAX; ++m) {
        if (m >= cnt) break;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const double v = wave_sum(acc[m][q]);
            if (lane == 0) s_red[wave][m * 4 + q] = v;
        }
    }
    __syncthreads();
    if ((int)threadIdx.x < cnt * 4) {
        const int m = threadIdx.x >> 2, q = threadIdx.x & 3;
        double sum = 0.0;
        for (int wv = 0; wv < PREP_THREADS / 64; ++wv) sum += s_red[wv][m * 4 + q];
        jobs[members[G.first + m]].scalars[q] = sum;
    }
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
// at most 10001 nodes.  One thread per job; node count written to the job.  A
// job flagged DWBA_PREP_NODES_SHARED reads a buffer another job of the launch
// fills with the same stepping: it steps only to count its nodes.
__global__ void coulomb_nodes_kernel(dwba_prep_job* __restrict__ jobs, int n_jobs, const double* __restrict__ par) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_jobs) return;
    dwba_prep_job& J = jobs[k];
    if ((J.mode & DWBA_PREP_MODE_MASK) != DWBA_PREP_FILON || J.phase_nodes == nullptr) return;
    const double a = par[4 * k + 0], b = par[4 * k + 1], kt = par[4 * k + 2], eta = par[4 * k + 3];
    const double dphi = J.p_step;  // phase increment for Coulomb jobs
    double* nodes = (J.mode & DWBA_PREP_NODES_SHARED) ? nullptr : const_cast<double*>(J.phase_nodes);
    int n = 0;
    if (nodes) nodes[n] = a;
    ++n;
    double r = a;
    const double cap = (b - a) / 10;
    while (r < b) {
        const double pp = kt + eta / fmax(r, 1e-10);
        double dr = dphi / fabs(pp);
        dr = fmin(dr, cap);
        dr = fmax(dr, 1e-6);
        double rn = r + dr;
        if (rn >= b) rn = b;
        if (nodes) nodes[n] = rn;
        ++n;
        r = rn;
        if (n > 10000) break;
    }
    J.n_pnodes = n;
}

}  // namespace

// CC table with nodes ascending in x (the FILON gather relies on it; the sum order is free).
bool make_cc(const double* cc_x, const double* cc_w, int n_cc, CCTable* cc) {
    if (n_cc < 2 || n_cc > DWBA_MAX_CC) return false;
    int ord[DWBA_MAX_CC];
    for (int k = 0; k < n_cc; ++k) ord[k] = k;
    std::sort(ord, ord + n_cc, [&](int p, int q) { return cc_x[p] < cc_x[q]; });
    for (int k = 0; k < n_cc; ++k) { cc->x[k] = cc_x[ord[k]]; cc->w[k] = cc_w[ord[k]]; }
    cc->n = n_cc;
    return true;
}

int launch_prep(int n_jobs, dwba_prep_job* d_jobs, const double* d_coulomb_par, int any_coulomb,
                const double* cc_x, const double* cc_w, int n_cc, hipStream_t s) {
    CCTable cc{};
    if (!make_cc(cc_x, cc_w, n_cc, &cc))
        return set_error(DWBA_ERR_ARG, "prep: CC node count %d outside [2, %d]", n_cc, DWBA_MAX_CC);
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

extern "C" int32_t dwba_filon_cells_f64(int32_t n_groups, const dwba_filon_group* d_groups,
                                       const dwba_prep_job* d_jobs, const double* h_cc_x, const double* h_cc_w,
                                       int32_t n_cc, void* stream) {
    DWBA_REQUIRE(dwba::initialised(), "dwba_filon_cells_f64: call dwba_init first");
    DWBA_REQUIRE(n_groups > 0 && d_groups && d_jobs && h_cc_x && h_cc_w, "dwba_filon_cells_f64: bad arguments");
    dwba::CCTable cc{};
    DWBA_REQUIRE(dwba::make_cc(h_cc_x, h_cc_w, n_cc, &cc), "dwba_filon_cells_f64: CC node count %d outside [2, %d]",
                 n_cc, DWBA_MAX_CC);
    hipLaunchKernelGGL(dwba::filon_cells_kernel, dim3(n_groups), dim3(dwba::PREP_THREADS), 0,
                       reinterpret_cast<hipStream_t>(stream), d_groups, d_jobs, cc);
    DWBA_LAUNCH_CHECK("filon_cells_kernel");
    return DWBA_OK;
}

extern "C" int32_t dwba_filon_fused_f64(int32_t n_slices, const dwba_filon_group* d_slices,
                                       const dwba_prep_job* d_jobs, const int32_t* d_members, void* stream) {
    DWBA_REQUIRE(dwba::initialised(), "dwba_filon_fused_f64: call dwba_init first");
    DWBA_REQUIRE(n_slices > 0 && d_slices && d_jobs && d_members, "dwba_filon_fused_f64: bad arguments (n_slices=%d)",
                 n_slices);
    hipLaunchKernelGGL(dwba::filon_fused_kernel, dim3(n_slices), dim3(dwba::PREP_THREADS), 0,
                       reinterpret_cast<hipStream_t>(stream), d_slices, d_jobs, d_members);
    DWBA_LAUNCH_CHECK("filon_fused_kernel");
    return DWBA_OK;
}
