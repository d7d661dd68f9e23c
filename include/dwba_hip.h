/*
 * dwba_hip.h -- C ABI of the MI355X (gfx950) radial multipole integral library.
 *
 * I_L = sum_ij alpha_i K_L(r_i, r_j) beta_j,  K_L = r_<^L / r_>^(L+1)
 * (diagonal (1-1e-12)^L / r_i), for L = 0..L_max, direct and exchange.
 *
 * Conventions
 *  - Every pointer argument named d_* is DEVICE memory owned by the caller;
 *    h_* is host memory.  The library allocates nothing behind the caller's
 *    back except the per-process state created by dwba_init and the one
 *    device allocation a scan handle owns (dwba_scan_plan .. dwba_scan_destroy).
 *  - `stream` is a hipStream_t (NULL = default stream).  Calls only enqueue
 *    work; results are valid after the caller synchronises the stream.
 *  - Every function returns DWBA_OK (0) or an error code; the message of the
 *    last error on the calling thread is returned by dwba_last_error().
 *    No C++ exception crosses this boundary.
 *
 * Reference interfaces replaced (StaryPasterz/DWBA, pure Python + CuPy):
 *  - dwba_matrix_elements.py:1563-2549 radial_ME_all_L_gpu, whose O(N^2) work
 *    is the per-L kernel build + GEMV at :2318-2515 (cp.dot at :2350, :2357,
 *    :2374, :2393, :2476-2477, :2508-2509); the CPU semantics it must match
 *    are radial_ME_all_L, dwba_matrix_elements.py:780-1300.
 */
#ifndef DWBA_HIP_H
#define DWBA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    DWBA_OK = 0,
    DWBA_ERR_ARG = 1,        /* invalid argument (shape, pointer, range) */
    DWBA_ERR_HIP = 2,        /* HIP runtime error */
    DWBA_ERR_NOT_INIT = 3,   /* dwba_init was not called */
    DWBA_ERR_WORKSPACE = 4   /* workspace too small */
};

/* Library version, 100*major + minor. */
int32_t dwba_version(void);

/* Selects `device` for the calling thread and records its properties.
 * Replaces the CuPy runtime probe check_cupy_runtime()
 * (dwba_matrix_elements.py:178-193). */
int32_t dwba_init(int32_t device);
int32_t dwba_shutdown(void);
const char* dwba_last_error(void);

/* ------------------------------------------------------------------ */
/* Plain FP64 contraction, all L, direct + exchange, one channel pair.  */
/* ------------------------------------------------------------------ */
/* Computes, for L = 0..L_max,
 *   out[L]            = alpha_dir^T K_L beta_dir
 *   out[L_max+1 + L]  = alpha_ex^T  K_L beta_ex
 * over rows i < n_rows and columns j < n_cols of the grid d_r (n_cols >=
 * n_rows points, strictly increasing, > 0).  alpha_* must be zero at
 * i >= n_rows (they are read only there); beta_* are read on [0, n_cols).
 * K_L is generated on chip, never stored.  FP64 throughout.
 *
 * Replaces the per-L `np.dot(kernel_L, rho2_w)` / `np.dot(rho1_w, int_r2)`
 * pair of radial_ME_all_L (dwba_matrix_elements.py:1243-1245, :1288-1289)
 * and of oscillatory_kernel_integral_2d (oscillatory_integrals.py:2571,
 * :2756), and the CuPy GEMVs of radial_ME_all_L_gpu
 * (dwba_matrix_elements.py:2476-2477, :2508-2509).
 * L_max in [0, 63].  workspace >= dwba_il_plain_f64_workspace_bytes(). */
size_t dwba_il_plain_f64_workspace_bytes(int32_t n_rows, int32_t n_cols, int32_t L_max);
int32_t dwba_il_plain_f64(const double* d_r, int32_t n_rows, int32_t n_cols,
                          const double* d_alpha_dir, const double* d_beta_dir,
                          const double* d_alpha_ex, const double* d_beta_ex,
                          int32_t L_max, double* d_out,
                          void* d_workspace, size_t workspace_bytes, void* stream);

/* The same as dwba_il_plain_f64 with no workspace argument and no allocation:
 * no counterpart in the reference.  It is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, hosts
 * that forbid hipMalloc).  The kernels' partial sums go to a fixed
 * library-owned device array of DWBA_NOALLOC_SCRATCH_BYTES per GPU (held
 * until the process ends; its address on each device is taken by dwba_init,
 * which must have been called for the current device).  All calls on one
 * GPU share that array, so calls that overlap in time on one GPU (two
 * streams, two threads) race: serialise them.  Returns DWBA_ERR_WORKSPACE if
 * the launch would need more scratch than that (more than 512 CUs).  Same
 * arguments and bit-identical results as dwba_il_plain_f64. */
#define DWBA_NOALLOC_SCRATCH_BYTES (8u << 20)
int32_t dwba_il_plain_f64_noalloc(const double* d_r, int32_t n_rows, int32_t n_cols,
                                  const double* d_alpha_dir, const double* d_beta_dir,
                                  const double* d_alpha_ex, const double* d_beta_ex,
                                  int32_t L_max, double* d_out, void* stream);

/* Batched form: n_jobs independent channel pairs (possibly on different
 * grids) in ONE launch.  d_jobs is a DEVICE array of n_jobs descriptors whose
 * pointers are device pointers; h_n_rows / h_n_cols are HOST copies of the
 * extents (used to size the launch).  Job k writes 2*(L_max+1) doubles at
 * d_out + k*2*(L_max+1), laid out as in dwba_il_plain_f64.  No host-device
 * copy is made, so the call can be captured in a HIP graph.
 * Replaces the per-pair sequential calls of driver.py:907-1029 /
 * ionization.py:991-1002 into radial_ME_all_L_gpu. */
typedef struct dwba_plain_job {
    const double* r;          /* grid, n_cols points */
    const double* alpha_dir;  /* n_rows (zero beyond) */
    const double* beta_dir;   /* n_cols */
    const double* alpha_ex;
    const double* beta_ex;
    int32_t n_rows;
    int32_t n_cols;
} dwba_plain_job;

size_t dwba_il_plain_f64_batch_workspace_bytes(int32_t n_jobs, const int32_t* h_n_rows,
                                               const int32_t* h_n_cols, int32_t L_max);
int32_t dwba_il_plain_f64_batch(int32_t n_jobs, const dwba_plain_job* d_jobs,
                                const int32_t* h_n_rows, const int32_t* h_n_cols,
                                int32_t L_max, double* d_out,
                                void* d_workspace, size_t workspace_bytes, void* stream);

/* O(N)-per-L form of dwba_il_plain_f64_batch: same jobs, same output layout,
 * same FP64 values up to rounding order.  K_L is semiseparable on an
 * increasing grid (K_L(i,j) = (r_j/r_i)^L / r_i for j < i), so
 *   I_L = sum_i (alpha_i A_i + beta_i Z_i + (1-1e-12)^L alpha_i beta_i) / r_i,
 *   A_{i+1} = (r_i/r_{i+1})^L (A_i + beta_i),  Z_{i+1} = (r_i/r_{i+1})^L (Z_i + alpha_i),
 * evaluated as wavefront-parallel affine scans.  Replaces the same reference
 * code as dwba_il_plain_f64 (dwba_matrix_elements.py:1057-1080, :1113-1116,
 * :1243-1245, :1281-1289).  Requires 0 <= n_rows <= n_cols and d_r strictly
 * increasing and > 0.  A job whose alpha_dir is NULL has its direct family
 * skipped (the direct outputs are left as they are; see
 * dwba_semisep_gemv_f64).  Tuning (results agree to rounding for every
 * choice): seg = grid points per lane (0 = chosen from the work size; the
 * workspace size depends on it), lb = multipoles per register block (4 or 8,
 * 0 = 8). */
size_t dwba_il_semisep_f64_workspace_bytes(int32_t n_jobs, const int32_t* h_n_rows, const int32_t* h_n_cols,
                                           int32_t L_max, int32_t seg);
int32_t dwba_il_semisep_f64(int32_t n_jobs, const dwba_plain_job* d_jobs, const int32_t* h_n_rows,
                            const int32_t* h_n_cols, int32_t L_max, int32_t seg, int32_t lb, double* d_out,
                            void* d_workspace, size_t workspace_bytes, void* stream);

/* Lane-serial form of dwba_il_semisep_f64: the same recurrences, one lane per
 * (job, family, block of 8 multipoles) over the job's whole grid, so no
 * workspace and no join; same jobs, output layout and NULL-alpha_dir rule,
 * same values up to rounding order.  Fewer operations per grid point, less
 * parallelism: the better choice when jobs are many (the autotuner decides).
 * lb = multipoles per lane (4 or 8, 0 = 8): 4 doubles the lanes.
 * Replaces the same reference code as dwba_il_semisep_f64. */
int32_t dwba_il_semisep_lanes_f64(int32_t n_jobs, const dwba_plain_job* d_jobs, int32_t L_max, int32_t lb,
                                  double* d_out, void* stream);

/* Table-driven form of dwba_il_semisep_lanes_f64 (same jobs, output and FP64
 * operations, so bit-identical results): the grid factors 1/r_i, q_i and
 * q_i^(lb*b) are read from a per-grid table instead of being recomputed by
 * every lane.  dwba_semisep_lane_tables_f64 writes, for each grid g with
 * n points, tab[i*S + 0] = 1/r_i, tab[i*S + 1] = r_i * (1/r_{i+1}) (index
 * clamped to n-1), tab[i*S + 2 + b] = (q_i^lb)^b for b < S - 2, where
 * S = dwba_semisep_lane_table_stride(L_max, lb).  d_tabs[k] (device array of
 * device pointers) is the table of job k's grid, covering >= n_cols points.
 * d_out_index (device, may be NULL = identity): job k writes output row
 * d_out_index[k], so jobs may be given in any order (longest first balances
 * the wavefronts).  Replaces the same reference code as dwba_il_semisep_f64. */
typedef struct dwba_lane_grid {
    const double* r;          /* n points, increasing */
    double* tab;              /* n * stride doubles */
    int32_t n;
    int32_t pad_;
} dwba_lane_grid;
int32_t dwba_semisep_lane_table_stride(int32_t L_max, int32_t lb);
int32_t dwba_semisep_lane_tables_f64(int32_t n_grids, const dwba_lane_grid* d_grids, int32_t max_n, int32_t L_max,
                                     int32_t lb, void* stream);
int32_t dwba_il_semisep_lanes_tab_f64(int32_t n_jobs, const dwba_plain_job* d_jobs, const double* const* d_tabs,
                                      const int32_t* d_out_index, int32_t L_max, int32_t lb, double* d_out,
                                      void* stream);

/* The same with each (job, family, block) range split between two
 * wavefronts (the second starts from zero carries, accumulates the
 * propagation sums of the first's carries and joins it through LDS): a
 * shorter dependency chain per wavefront for 1.3x the operations; same
 * arguments and tables, same values up to rounding order.  waves_per_simd
 * (0 = 2, or 3): the occupancy the kernel is compiled for (3 caps it at
 * 168 VGPRs).  Replaces the same reference code as dwba_il_semisep_f64. */
int32_t dwba_il_semisep_lanes_split_f64(int32_t n_jobs, const dwba_plain_job* d_jobs, const double* const* d_tabs,
                                        const int32_t* d_out_index, int32_t L_max, int32_t lb,
                                        int32_t waves_per_simd, double* d_out, void* stream);

/* LDS-staged form of the split kernel (lb = multipoles per lane, 4 or 8,
 * 0 = 4; tables of stride dwba_semisep_lane_table_stride(L_max, lb)): the
 * G = dwba_semisep_lane_stage_group(L_max, lb) streams of a group (64 / the
 * lanes per stream: 8 for lb 4 and L_max 16..31) run at the same grid
 * positions, split at one point for the group, and every wavefront copies its
 * streams' alpha / beta and the group's grid-table rows into LDS chunks of 16
 * points ahead of use.  Streams are as in the split kernel (the direct family
 * of jobs [0, n_jobs), then the exchange family); the streams of group g are
 * [g G, g G + G).  Precondition: within a group, every stream with
 * n_cols > 0 has the same d_tabs pointer (pad each grid's jobs to a multiple
 * of G, and n_jobs with them, by jobs with n_rows = n_cols = 0); a group that
 * breaks it writes NaN for all of its streams.  d_out_index (may be NULL =
 * identity): job k writes output row d_out_index[k], none if negative.
 * dwba_semisep_lane_stage_group returns -1 where the kernel does not apply
 * (lb 4 needs L_max >= 4, lb 8 L_max >= 16).  Same values up to rounding
 * order.  Replaces the same reference code as dwba_il_semisep_f64. */
int32_t dwba_semisep_lane_stage_group(int32_t L_max, int32_t lb);
int32_t dwba_il_semisep_lanes_stage_f64(int32_t n_jobs, const dwba_plain_job* d_jobs, const double* const* d_tabs,
                                        const int32_t* d_out_index, int32_t L_max, int32_t lb, double* d_out,
                                        void* stream);

/* ------------------------------------------------------------------ */
/* Full-Split preparation: alpha / beta vectors of a channel pair.       */
/* ------------------------------------------------------------------ */
#define DWBA_MAX_CC 16

enum {
    DWBA_PREP_PLAIN = 0,     /* oscillatory quadrature off or k_i+k_f <= k_threshold */
    DWBA_PREP_FILON = 1,     /* Filon / Clenshaw-Curtis direct + exchange            */
    DWBA_PREP_STANDARD = 2,  /* Filon branch fallback to weighted dot products       */
    DWBA_PREP_ZERO = 3       /* idx_limit < 2 in the Filon branch: integral is 0     */
};
/* Flag or-ed into dwba_prep_job.mode (FILON mode): skip clearing alpha_dir,
 * alpha_ex and beta_ex and write only the points the Clenshaw-Curtis nodes
 * touch.  For callers that keep one zero-initialised set of vectors per pair
 * and prepare the same pair again (a captured scan): the untouched points are
 * the same on every call and keep their zeros. */
#define DWBA_PREP_NO_ZERO_FILL 0x10
#define DWBA_PREP_FILON_GROUPED 0x20   /* FILON: nothing in this launch (Coulomb nodes aside); alpha/beta,
                                          beta_dir and the scalars come from dwba_filon_cells_f64 +
                                          dwba_filon_fused_f64 */
#define DWBA_PREP_NODES_SHARED 0x40    /* FILON with Coulomb stepping: phase_nodes is the buffer of another job of
                                          the same launch with the same d_coulomb_par entry and p_step (pairs of one
                                          Filon group); that job fills it, this one only counts its nodes */
#define DWBA_PREP_FINITE_TAIL 0x80     /* FILON grouped: chi_i, chi_f, u_f, u_i and V_diff are finite on [N1, N), so
                                          the L = 0 sums gain exact zeros there (w_lim = 0) and dwba_filon_fused_f64
                                          stops at N1 for a slice whose members all carry it and write no beta_dir */
#define DWBA_PREP_MODE_MASK 0x0f

/* One channel pair.  Device pointers; outputs are the alpha and beta vectors (N doubles
 * each) and scalars[4] = {sum rho2_dir_w, rho1_dir_w.V_diff, sum rho2_ex_w,
 * rho1_ex_w.V_diff} (the L = 0 correction factors).  Phase intervals: either
 * uniform (phase_nodes == NULL: node m = m*p_step + p_start, the last node
 * p_end, n_pnodes nodes -- numpy linspace) or an explicit ascending array of
 * n_pnodes nodes; for Coulomb stepping the array is filled on the device
 * (see dwba_fullsplit_prep_f64) and p_step holds the phase increment. */
typedef struct dwba_prep_job {
    const double* r;          /* grid, N points */
    const double* logr;       /* log(max(r, 1e-300)), N points */
    const double* w;          /* Simpson weights, N points */
    const double* u_i;        /* bound_i.u_of_r */
    const double* u_f;        /* bound_f.u_of_r (continuum for ionization) */
    const double* chi_i;      /* cont_i.chi_of_r */
    const double* chi_f;      /* cont_f.chi_of_r */
    const double* V_diff;     /* V_core - U_i, or NULL */
    const double* phase_nodes;
    double* alpha_dir;
    double* beta_dir;         /* may be NULL in FILON / STANDARD mode: beta_dir = (u_f u_i) w is then not
                                 written (pairs of one energy that share it need it once) */
    double* alpha_ex;
    double* beta_ex;
    double* scalars;
    double p_start, p_step, p_end;
    int32_t N, N1;            /* grid size, split index idx_limit */
    int32_t n_pnodes;
    int32_t mode;             /* DWBA_PREP_* */
    int32_t log1, log2;       /* r[:N1] / r[:N] log-spaced (5e-3 test) */
} dwba_prep_job;

/* Builds alpha/beta for n_jobs pairs in one launch (one workgroup per pair).
 * d_jobs: DEVICE array.  Jobs in FILON mode with a non-NULL phase_nodes
 * buffer (>= 10002 doubles) get their nodes generated on the device by the
 * Coulomb stepping of generate_phase_nodes when any_coulomb != 0, from
 * d_coulomb_par[4*k..4*k+3] = {r_start, r_end, k_total, eta_total}.
 * h_cc_x / h_cc_w: the n_cc reference Clenshaw-Curtis nodes/weights on
 * [-1, 1] (host).
 * Replaces oscillatory_kernel_integral_2d's node construction and
 * interpolations (oscillatory_integrals.py:2578-2644, :2667-2770) and the
 * density preparation of radial_ME_all_L (dwba_matrix_elements.py:1025-1049). */
int32_t dwba_fullsplit_prep_f64(int32_t n_jobs, dwba_prep_job* d_jobs, const double* d_coulomb_par,
                                int32_t any_coulomb, const double* h_cc_x, const double* h_cc_w,
                                int32_t n_cc, void* stream);

/* Filon cell coefficients shared by a group of FILON jobs with the same grid,
 * split index N1, phase layout and interpolation kinds (pairs of one energy
 * that differ only in their densities).  For cell j < N1 - 1 the CC nodes in
 * [r_j, r_j+1) contribute to alpha at points j, j+1 through the 2x2 sums
 * c00, c01, c11 (of W (1-t)^2, W t (1-t), W t^2) and to beta_ex through
 * b00, b01, b10, b11 (of W wl (1-t), W wl t, W wr (1-t), W wr t), exactly as
 * dwba_fullsplit_prep_f64 forms them per pair.  coef holds these 7 rows at
 * stride ld_coef (>= N1); rep is the index in d_jobs of one member (its
 * phase nodes, for Coulomb stepping, must already be generated, i.e. the
 * prep launch with DWBA_PREP_FILON_GROUPED jobs runs first); members[first ..
 * first + count) are the jobs the apply step writes.  Replaces the per-pair
 * node interpolations of oscillatory_kernel_integral_2d
 * (oscillatory_integrals.py:2578-2644, :2667-2770). */
typedef struct dwba_filon_group {
    double* coef;
    int32_t ld_coef;
    int32_t rep;
    int32_t first;
    int32_t count;
} dwba_filon_group;
int32_t dwba_filon_cells_f64(int32_t n_groups, const dwba_filon_group* d_groups, const dwba_prep_job* d_jobs,
                             const double* h_cc_x, const double* h_cc_w, int32_t n_cc, void* stream);
/* Everything dwba_fullsplit_prep_f64 writes for the member jobs of each
 * slice (<= 4 members of one group: entries with the same coef and rep and
 * disjoint member ranges), from the group's cells: alpha_dir, alpha_ex,
 * beta_ex on [0, N1) (point j takes cell j's part c00 rho_j + c01 rho_j+1 and
 * cell j-1's part c01 rho_j-1 + c11 rho_j), beta_dir = (u_f u_i) w on [0, N)
 * where the job has one, and the four L = 0 correction scalars; the jobs'
 * densities are read once.  The members share the representative's grid, N,
 * N1, u_i and V_diff.  With dwba_filon_cells_f64 it replaces the density
 * interpolations at the nodes (oscillatory_integrals.py:2578-2644,
 * :2667-2770) and the density preparation of radial_ME_all_L
 * (dwba_matrix_elements.py:1025-1049).  Slices with count 0 are skipped.
 * Placement: the slices are cut into 8 consecutive runs of
 * ceil(n_slices/8) (the first n_slices % 8 runs one longer), and run q goes
 * to XCD q, in order; listing each energy's slices inside one run keeps its
 * shared arrays and cell tables in one L2 (results do not depend on it). */
int32_t dwba_filon_fused_f64(int32_t n_slices, const dwba_filon_group* d_slices, const dwba_prep_job* d_jobs,
                             const int32_t* d_members, void* stream);

/* Multipole moments M_L = sum_j ((w_j r_j^L) u_f,j) u_i,j for L = 0..L_max of
 * n_sets density pairs (device pointer arrays d_r/d_w/d_u_i/d_u_f of length
 * n_sets, extents h_n on the host), written to d_out[set*(L_max+1) + L].
 * Replaces `np.sum(w * (r ** L) * u_f * u_i)` (dwba_matrix_elements.py:1167,
 * :1216) and its GPU counterpart (:2150-2170). */
int32_t dwba_moments_f64(int32_t n_sets, const double* const* d_r, const double* const* d_w,
                         const double* const* d_u_i, const double* const* d_u_f, const int32_t* d_n,
                         int32_t L_max, double* d_out, void* stream);

/* ------------------------------------------------------------------ */
/* Outer oscillatory tail of the direct integral (FP64).                 */
/* ------------------------------------------------------------------ */
enum {
    DWBA_TAIL_NONE = 0,    /* no tail (L = 0, ionization, idx_limit >= N-10, ...) */
    DWBA_TAIL_OUTER = 1,   /* Filon/Levin on [r_m, r_max] + analytic tail at r_max */
    DWBA_TAIL_LEGACY = 2   /* analytic tail at r_m only (method "legacy")          */
};

/* One channel pair: incident wave a = (k, l, delta, eta, sigma)_i, scattered
 * wave b = (...)_f, split radius r_m = r[idx_limit-1], r_max = r[N-1],
 * moments[L] (device, L = 0..L_max), out[L] (device).  out[0] = 0 and
 * out[L] = 0 where |moments[L]| <= 1e-12. */
typedef struct dwba_tail_job {
    double k_a, delta_a, eta_a, sigma_a;
    double k_b, delta_b, eta_b, sigma_b;
    int32_t l_a, l_b;
    double r_m, r_max;
    const double* moments;
    double* out;
    int32_t L_max;
    int32_t kind;             /* DWBA_TAIL_* */
} dwba_tail_job;

/* One wavefront per (pair, sign of the sinA*sinB split) in one launch; each
 * wavefront evaluates the phases and Filon coefficients once per segment and
 * sweeps all L.  max_filon_segments / max_effective_dphi are
 * DWBA_MAX_FILON_SEGMENTS / DWBA_MAX_EFFECTIVE_DPHI; h_levin_D (8x8
 * row-major) and h_levin_x (8) are the Chebyshev differentiation matrix and
 * the ascending Lobatto points of the Levin rule.  L_max in [0, 63];
 * workspace >= dwba_outer_tail_workspace_bytes(n_jobs, L_max).
 * Replaces dwba_outer_integral_1d + _analytical_multipole_tail
 * (oscillatory_integrals.py:459-535, :1024-1451, :2064-2184) as called at
 * dwba_matrix_elements.py:1165-1192 and :1257-1268. */
size_t dwba_outer_tail_workspace_bytes(int32_t n_jobs, int32_t L_max);
int32_t dwba_outer_tail_f64(int32_t n_jobs, const dwba_tail_job* d_jobs, int32_t L_max,
                            int32_t max_filon_segments, double max_effective_dphi,
                            const double* h_levin_D, const double* h_levin_x, void* d_workspace,
                            size_t workspace_bytes, void* stream);

/* ------------------------------------------------------------------ */
/* Batched scan: all channel pairs of all energies in one launch each.    */
/* ------------------------------------------------------------------ */
/* Exchange family on MFMA.  One descriptor per energy; pairs are columns.
 * B: beta_ex of pair p at B + p*ld (FP32, zero beyond its support and for
 * ld > j >= N); A: alpha_ex (FP64, same layout, zero beyond N1_p).  Pairs
 * are grouped in column blocks of 128 (ideally sorted by support);
 * col_extent[cb] is the largest beta support in block cb and row_blocks[cb]
 * the number of 256-row blocks covering the largest alpha support.  Partial
 * sums of (L, cb, rb) live at partial + part_off + ((L*n_cb + cb)*n_rb + rb)*128. */
typedef struct dwba_xgemm_energy {
    const double* logr;
    const float* B;           /* FP32 beta (dwba_xgemm_f32) */
    const uint16_t* Bh;       /* bf16 hi / lo parts of beta (dwba_xgemm_bf16x3) */
    const uint16_t* Bl;
    const double* A;
    const int32_t* col_extent;
    const int32_t* row_blocks;
    int64_t part_off;
    int32_t N, P, ld, n_cb, n_rb;
    int32_t pair_base;        /* index of this energy's first pair in d_out */
    int32_t out_slot;         /* 0: direct, 1: exchange half of d_out */
    int32_t pad_;
} dwba_xgemm_energy;

/* hi[i] = bf16(x[i]), lo[i] = bf16(x[i] - hi[i]) (round to nearest even):
 * the two-term bf16 split of the beta columns for dwba_xgemm_bf16x3. */
int32_t dwba_split_bf16(const double* d_x, uint16_t* d_hi, uint16_t* d_lo, int64_t n, void* stream);

/* Split-BF16x3 variant of dwba_xgemm_f32: K and beta split into bf16 hi + lo,
 * T += Khi Bhi + Khi Blo + Klo Bhi on v_mfma_f32_32x32x16_bf16, FP32
 * accumulation, FP64 epilogue.  Reads E.Bh / E.Bl instead of E.B. */
int32_t dwba_xgemm_bf16x3(int32_t n_energies, const dwba_xgemm_energy* d_energies, int32_t n_items,
                          const int32_t* d_items, int32_t L_max, double* d_partial, int32_t* d_counter,
                          int32_t max_cb, double* d_out, int32_t n_workgroups, void* stream);

/* y[i] = (float)x[i], i < n (rounding of the beta columns for the FP32 MFMA path). */
int32_t dwba_cast_f64_f32(const double* d_x, float* d_y, int64_t n, void* stream);

/* Doubles of partial storage one energy needs. */
size_t dwba_xgemm_partial_doubles(int32_t n_rb, int32_t n_cb, int32_t L_max);

/* d_items: n_items work items (int32 x4: energy, L, cb, rb) in the order they
 * should start (longest first); d_counter: one int32 of scratch (reset by the
 * call with a kernel, so the call can be captured in a HIP graph).  Writes
 * d_out[(pair_base+p)*2*(L_max+1) + out_slot*(L_max+1) + L] for every pair p.
 * n_workgroups <= 0 picks one persistent workgroup per CU. */
int32_t dwba_xgemm_f32(int32_t n_energies, const dwba_xgemm_energy* d_energies, int32_t n_items,
                       const int32_t* d_items, int32_t L_max, double* d_partial, int32_t* d_counter,
                       int32_t max_cb, double* d_out, int32_t n_workgroups, void* stream);

/* Block form of the stacked MFMA contraction (same results as dwba_xgemm_f32
 * / _bf16x3 up to their rounding, same reference code replaced:
 * dwba_matrix_elements.py:1281-1289, :2508-2509).  K_L between two different
 * blocks of 8 grid points is rank one (r_j^L r_i^-(L+1)), so the contraction is
 * per block: four point moments (32 multipoles x 8 points, MFMA), the block's
 * own 36 (i <= j) entries (32 x 36, MFMA) and an FP64 scan over blocks.
 * Operands FP32 (v_mfma_f32_16x16x4_f32) or split-BF16x3
 * (v_mfma_f32_16x16x16_bf16, hi*hi + hi*lo + lo*hi); FP32 accumulation within
 * a block, FP64 between blocks.  L_max <= 31. */
enum {
    DWBA_XBLOCK_F32 = 0,
    DWBA_XBLOCK_BF16X3 = 1,
};
typedef struct dwba_xblock_grid {
    const double* r;          /* N radii, strictly increasing, > 0 */
    void* tab;                /* dwba_xblock_table_bytes(N, mode) bytes, written by dwba_xblock_tables */
    int32_t N;
    int32_t pad_;
} dwba_xblock_grid;
size_t dwba_xblock_table_bytes(int32_t N, int32_t mode);
/* The per-block factor tables of every grid (depend on the grid only). */
int32_t dwba_xblock_tables(int32_t n_grids, const dwba_xblock_grid* d_grids, int32_t max_n, int32_t mode,
                           void* stream);
/* Up to 64 consecutive pairs of one stacked matrix: alpha of pair c at
 * A + c*ld, beta at B + c*ld, both zero beyond their supports;
 * n_blocks = ceil(extent / 8) with extent >= every alpha and beta support of
 * the group and 8 n_blocks <= ld; range r (one workgroup) is blocks
 * [r blocks_per_range, (r + 1) blocks_per_range) of them, n_ranges =
 * ceil(n_blocks / blocks_per_range), and its partials sit at range0 + r of
 * d_partial.  Writes d_out[(out_row0 + c)*2*(L_max+1) + slot*(L_max+1) + L]. */
typedef struct dwba_xblock_group {
    const double* A;
    const double* B;
    const void* tab;          /* the grid's table */
    int32_t ld;
    int32_t n_pairs;          /* 1..64 */
    int32_t n_blocks;
    int32_t n_ranges;
    int32_t range0;
    int32_t out_row0;
    int32_t slot;             /* 0: direct, 1: exchange */
    int32_t blocks_per_range;
} dwba_xblock_group;
/* d_partial: dwba_xblock_partial_doubles(total ranges) doubles of scratch. */
size_t dwba_xblock_partial_doubles(int32_t n_ranges_total);
/* d_items: n_items int32 pairs (group, range) in dispatch order (workgroup i
 * runs item i; a group index < 0 is an empty item); every range of every
 * group appears once. */
int32_t dwba_xblock_contract(int32_t mode, int32_t n_groups, const dwba_xblock_group* d_groups, int32_t n_items,
                             const int32_t* d_items, int32_t L_max, double* d_partial, double* d_out,
                             void* stream);

/* Direct family (FP64 VALU): t[L*ld + i] = sum_j K_L(i, j) beta[j] for
 * i < rows, j < N, L = 0..L_max, for each energy (one shared beta per
 * energy).  max_rows = max over energies of rows. */
typedef struct dwba_dgemv_energy {
    const double* r;
    const double* beta;
    double* t;
    int32_t N, rows, ld, L_max;
} dwba_dgemv_energy;
int32_t dwba_dgemv_f64(int32_t n_energies, const dwba_dgemv_energy* d_energies, int32_t max_rows,
                       int32_t L_max, void* stream);

/* Same output as dwba_dgemv_f64 in O(N) per (energy, L) by the
 * semiseparable recurrences (forward for j < i, backward for j > i):
 * t_L(i) = (A_i + (1-1e-12)^L beta_i) / r_i + B_i.  The direct family of an
 * excitation scan has one beta per energy (dwba_matrix_elements.py:1136-1160,
 * oscillatory_integrals.py:2571), so I_p(L) = alpha_p^T t_L (dwba_ddot_f64). */
int32_t dwba_semisep_gemv_f64(int32_t n_energies, const dwba_dgemv_energy* d_energies, int32_t L_max,
                              void* stream);

/* d_out[out_index*2*(L_max+1) + L] = sum_{i<n} alpha[i] t[L*ld + i], one
 * workgroup per pair. */
typedef struct dwba_ddot_pair {
    const double* alpha;
    const double* t;
    int32_t n, ld, out_index, pad_;
} dwba_ddot_pair;
int32_t dwba_ddot_f64(int32_t n_pairs, const dwba_ddot_pair* d_pairs, int32_t L_max, double* d_out,
                      void* stream);

/* The same dot products for pairs grouped by shared t: d_items holds n_items
 * int32 pairs {first, count} (count <= 64), each a run of consecutive
 * d_pairs entries with the same t and ld; alpha must be zero from n up to
 * the largest n of its item (it is read there).
 * Each t row is read once per item instead of once per pair.  L_max <= 31. */
int32_t dwba_ddot_tiled_f64(int32_t n_items, const int32_t* d_items, const dwba_ddot_pair* d_pairs, int32_t L_max,
                            double* d_out, void* stream);

/* The same dot products on the FP64 matrix cores (v_mfma_f64_16x16x4_f64):
 * items of count <= 16 pairs sharing t and ld; as for the tiled form, alpha
 * must be zero from n up to the largest n of its item (it is read there);
 * nothing at or beyond that largest n is read.  L_max <= 31.  Replaces, as
 * dwba_ddot_tiled_f64 does, the per-pair direct dot of the Filon direct
 * branch (oscillatory_integrals.py:2571-2644 via dwba_matrix_elements.py:1136-1160). */
int32_t dwba_ddot_mfma_f64(int32_t n_items, const int32_t* d_items, const dwba_ddot_pair* d_pairs, int32_t L_max,
                           double* d_out, void* stream);

/* The same FP64 matrix-core dots with the operands staged through LDS: the
 * items of dwba_ddot_mfma_f64 (count <= 16 pairs sharing t and ld), whose
 * four wavefronts copy their chunks of 16 points as whole row chunks into
 * LDS one chunk ahead of use; the same contract on alpha (zero from n up to
 * the item's largest n, nothing at or beyond that read) and t (rows of ld
 * doubles, ld >= that n).  L_max <= 31.  Replaces the same reference code as
 * dwba_ddot_mfma_f64. */
int32_t dwba_ddot_stage_f64(int32_t n_items, const int32_t* d_items, const dwba_ddot_pair* d_pairs, int32_t L_max,
                            double* d_out, void* stream);

/* Per-pair assembly of I_L in place in d_out (layout [pair][2][L_max+1]):
 * contractions zeroed when non-finite in the oscillatory branch, tail added
 * (DWBA_TAIL_OUTER before, DWBA_TAIL_LEGACY after the L = 0 correction),
 * L = 0 correction scalars[1]*scalars[0] / scalars[3]*scalars[2] when
 * |scalars[0|2]| > 1e-12 -- the order of radial_ME_all_L
 * (dwba_matrix_elements.py:1112-1298). */
typedef struct dwba_finalize_pair {
    const double* scalars;    /* 4 values from dwba_fullsplit_prep_f64 */
    const double* tails;      /* L_max+1 values from dwba_outer_tail_f64 (or NULL if none) */
    int32_t out_index, oscillatory, tail_kind, pad_;
} dwba_finalize_pair;
int32_t dwba_batch_finalize(int32_t n_pairs, const dwba_finalize_pair* d_pairs, int32_t L_max, double* d_out,
                            void* stream);

/* ------------------------------------------------------------------ */
/* A whole scan behind one handle: plan, launch, capture.                */
/* ------------------------------------------------------------------ */
/* The reference computes a scan's integrals one channel pair at a time:
 * radial_ME_all_L_gpu inside the l_i / l_f loops of compute_total_excitation_cs
 * (driver.py:907-1029), and for ionization one _compute_sdcs_at_energy call
 * per ejected energy (ionization.py:991-1002), each with its own pair loop
 * (:526-608).  dwba_scan_plan takes every pair of a scan at once, as HOST
 * arrays (the arrays those call sites hold), makes radial_ME_all_L's per-pair
 * decisions on the host (split index, Filon phase layout, tail kind;
 * dwba_matrix_elements.py:884-1108), and uploads grids, potentials and
 * waves into one device allocation together with every launch table.
 * dwba_scan_launch then enqueues the whole scan on a stream and allocates
 * nothing, so it can be captured (dwba_scan_capture, then dwba_scan_replay).
 * Results (device, valid once the stream reaches them): I_L of pair p at
 * dwba_scan_results(scan)[p*2*(L_max+1) + L] (direct) and
 * [p*2*(L_max+1) + L_max+1 + L] (exchange), pairs in the order given.
 * Arguments and array-length errors follow radial_ME_all_L_gpu
 * (dwba_matrix_elements.py:1637-1672: arrays longer than the grid are
 * truncated, shorter ones rejected).  The tail stage runs on a stream of the
 * scan's own, forked from and joined back into the launch stream.  One
 * scan's launches share its buffers: do not run two at once. */
typedef struct dwba_scan_energy {
    const double* r;          /* grid points (grid.r), n, increasing */
    const double* w_simpson;  /* integration weights (grid.w_simpson) */
    const double* w_trapz;    /* trapezoid weights (grid.w_trapz): the 99% extent of the bound states */
    const double* V_core;     /* core potential */
    const double* U_i;        /* distorting potential of the incident channel (U_i.U_of_r) */
    const double* U_f;        /* of the final channel (U_f_array), or NULL */
    const double* u_i;        /* initial bound state (bound_i.u_of_r) */
    int32_t n;                /* grid points, >= 2 */
    int32_t pad_;
} dwba_scan_energy;

typedef struct dwba_scan_wave {   /* one continuum wave (continuum.ContinuumWave) */
    const double* chi;        /* chi_of_r on the grid of the energy whose pairs use it */
    double k;                 /* k_au */
    double phase_shift, eta, sigma_l;
    int32_t l;
    int32_t idx_match;        /* match index; <= 0 or >= n: none */
    int32_t n;                /* points in chi (>= the grid's n) */
    int32_t pad_;
} dwba_scan_wave;

typedef struct dwba_scan_pair {
    const double* u_f;        /* bound_f.u_of_r (excitation) or the ejected wave's chi_of_r (ionization);
                                 pairs of an energy that pass the same pointer share its upload and moments */
    int32_t n_u_f;            /* points in u_f (>= the grid's n) */
    int32_t energy;           /* index into the energies */
    int32_t wave_i, wave_f;   /* indices into the waves: incident and scattered projectile */
    int32_t excitation;       /* 1: bound final state (outer tails apply); 0: ionization */
    int32_t pad_;
} dwba_scan_pair;

enum {                        /* oscillatory_method */
    DWBA_METHOD_LEGACY = 0,
    DWBA_METHOD_ADVANCED = 1,
    DWBA_METHOD_FULL_SPLIT = 2
};

typedef struct dwba_scan_options {   /* radial_ME_all_L's keyword arguments */
    int32_t L_max;                       /* 0..63 */
    int32_t use_oscillatory_quadrature;  /* reference default 1 */
    int32_t method;                      /* DWBA_METHOD_* (reference default "advanced") */
    int32_t CC_nodes;                    /* 2..16 (reference default 5) */
    double phase_increment;              /* reference default 1.5708 */
    double min_grid_fraction;            /* reference default 0.10 */
    double k_threshold;                  /* reference default 0.5 */
} dwba_scan_options;

typedef struct dwba_scan dwba_scan;   /* opaque */

/* Plans a scan (host decisions, one device allocation, uploads; synchronous).
 * On success *out_scan is the handle; on failure it is NULL. */
int32_t dwba_scan_plan(int32_t n_energies, const dwba_scan_energy* h_energies, int32_t n_waves,
                       const dwba_scan_wave* h_waves, int32_t n_pairs, const dwba_scan_pair* h_pairs,
                       const dwba_scan_options* h_options, dwba_scan** out_scan);
/* The per-pair decisions dwba_scan_plan makes, without a device (host only,
 * no dwba_init needed): for pair k, h_out[k] holds the split index N1
 * (idx_limit), the prep mode (DWBA_PREP_*), the tail kind (DWBA_TAIL_*), the
 * contraction extent (rows x cols), the log-spacing tests of r[:N1] and r[:N],
 * and the Filon phase layout (uniform: n_pnodes nodes from p_start to p_end
 * in steps p_step; Coulomb stepping: p_step is the phase increment and the
 * nodes are made on the device). */
typedef struct dwba_scan_pair_plan {
    double p_start, p_step, p_end;
    double r_match;           /* r[N1-1] */
    int32_t N1, mode, tail, rows, cols, log1, log2, n_pnodes, coulomb, oscillatory;
} dwba_scan_pair_plan;
int32_t dwba_scan_plan_pairs(int32_t n_energies, const dwba_scan_energy* h_energies, int32_t n_waves,
                             const dwba_scan_wave* h_waves, int32_t n_pairs, const dwba_scan_pair* h_pairs,
                             const dwba_scan_options* h_options, dwba_scan_pair_plan* h_out);
/* Enqueues the whole scan on stream (NULL = the null stream). */
int32_t dwba_scan_launch(dwba_scan* scan, void* stream);
/* Records dwba_scan_launch on stream (a created stream) into a HIP graph
 * owned by the scan, replacing an earlier one; nothing runs. */
int32_t dwba_scan_capture(dwba_scan* scan, void* stream);
/* Launches the captured graph on stream. */
int32_t dwba_scan_replay(dwba_scan* scan, void* stream);
/* Device pointer to the results (n_pairs * 2 * (L_max+1) doubles). */
const double* dwba_scan_results(const dwba_scan* scan);
/* Copies the results to host memory h_out (same layout) after the work
 * enqueued on stream so far; returns when they are there. */
int32_t dwba_scan_read_results(const dwba_scan* scan, double* h_out, void* stream);
/* h_info[0..n_info): pairs, L_max, device bytes, pairs with a tail, Filon
 * groups, energies with a shared direct beta, Coulomb node buffers, moment
 * sets (further entries 0). */
int32_t dwba_scan_info(const dwba_scan* scan, int64_t* h_info, int32_t n_info);
/* Waits for the device, then frees the scan (NULL is a no-op). */
int32_t dwba_scan_destroy(dwba_scan* scan);

/* ---- differentiation of the plain contraction -------------------------
 * Forward quantity (the plain branch of radial_ME_all_L,
 * dwba_matrix_elements.py:1243-1245 and :1288-1289, for P pairs sharing one
 * grid): I_L(p) = sum_ij w_i rho_a[p][i] K_L(i,j) w_j rho_b[p][j].
 * The reference has no gradient; these entry points give, for the loss
 * sum_{p,L} g[p][L] I_L(p),
 *   grad_a[p][i] = w_i sum_j Keff_p(i,j) w_j rho_b[p][j]
 *   grad_b[p][i] = w_i sum_j Keff_p(i,j) w_j rho_a[p][j],   Keff_p = sum_L g[p][L] K_L,
 * regenerating Keff element by element (nothing of size N x N is stored).
 * rho_a, rho_b, grad_a, grad_b: [n_pairs][ld] doubles (rows i >= n of the
 * outputs are written as 0); g: [n_pairs][L_max+1]; either output may be
 * NULL.  _f64 evaluates Keff and the products in FP64; _f32 in FP32 with
 * each 256-column chunk summed in FP32 and the chunks in FP64. */
int32_t dwba_il_backward_f64(const double* d_r, const double* d_w, int32_t n, int32_t n_pairs, int32_t ld,
                             const double* d_rho_a, const double* d_rho_b, const double* d_g, int32_t L_max,
                             double* d_grad_a, double* d_grad_b, void* stream);
int32_t dwba_il_backward_f32(const double* d_r, const double* d_w, int32_t n, int32_t n_pairs, int32_t ld,
                             const double* d_rho_a, const double* d_rho_b, const double* d_g, int32_t L_max,
                             double* d_grad_a, double* d_grad_b, void* stream);

/* The same gradients (FP64, same arguments and output layout as
 * dwba_il_backward_f64) in O(N) per pair and multipole: K_L x is formed by
 * the semiseparable recurrences of dwba_semisep_gemv_f64 for x = w rho_b[p]
 * and x = w rho_a[p], every L, then grad = w sum_L g[p][L] (K_L x).  Nothing of
 * size N x N is formed; the workspace (>= the _workspace_bytes value) holds
 * the weighted rows and the 2 n_pairs (L_max + 1) vectors K_L x.  The forward
 * quantity is the plain branch of radial_ME_all_L
 * (dwba_matrix_elements.py:1243-1245, :1288-1289); the reference has no
 * gradient. */
size_t dwba_il_backward_semisep_f64_workspace_bytes(int32_t n, int32_t n_pairs, int32_t L_max);
int32_t dwba_il_backward_semisep_f64(const double* d_r, const double* d_w, int32_t n, int32_t n_pairs, int32_t ld,
                                     const double* d_rho_a, const double* d_rho_b, const double* d_g, int32_t L_max,
                                     double* d_grad_a, double* d_grad_b, void* d_workspace, size_t workspace_bytes,
                                     void* stream);

/* ---- amplitude assembly -------------------------------------------------
 * The per-(M_i, M_f) sums that turn I_L into scattering amplitudes:
 * calculate_amplitude_contribution (dwba_coupling.py:285-504) accumulated
 * over the channel pairs of an energy as the driver does
 * (driver.py:1041-1052), calculate_ionization_coefficients
 * (dwba_coupling.py:507-634), dcs_dwba and integrate_dcs_over_angles
 * (sigma_total.py:147-217).  The angular coefficients are formed on the
 * host; a "type" is one (l_i, l_f, L_i, L_f) with its list of n_combos
 * projection pairs (M_i, M_f):
 *   d_coef  [type][combo][2][L_max+1]: Cd_L then Cx_L (0 where a term is cut);
 *   d_phase [type][combo][4]: complex constants (re, im) of the direct and
 *           exchange terms: i^(l_i+l_f) Y*_{l_i 0} and
 *           i^(l_i-l_f) (-1)^(L_f+M_f) (-1)^nu Y*_{l_i 0} (0 for a family
 *           whose Y_{l_f,m} does not exist).
 * d_integrals is [row][2][L_max+1] (direct, exchange) -- the layout of the
 * batched scan's output. */
typedef struct dwba_amp_pair {
    int32_t row;      /* the pair's row of d_integrals */
    int32_t type;     /* its coefficient table */
    int32_t l_f;      /* projectile final partial wave: F uses Y_{l_f, M_i - M_f}, G Y_{l_f, M_f - M_i} */
    int32_t pad_;
    double scale;     /* (2/pi)/(k_i k_f), times sqrt(2/pi)/k_ej for ionization */
} dwba_amp_pair;

/* d_ylm[(l*(2 m_max+1) + m + m_max)*n_theta + j] = Y_{l,m}(theta_j, phi = 0)
 * for l <= l_max, |m| <= m_max (Condon-Shortley phase, 0 where |m| > l), by the
 * normalised three-term recurrence in l.  The reference evaluates
 * scipy.special.sph_harm per call (dwba_coupling.py:430-470). */
int32_t dwba_ylm_f64(int32_t l_max, int32_t m_max, int32_t n_theta, const double* d_theta, double* d_ylm,
                     void* stream);

/* d_a[pair][combo][4] = scale * (phase_d S_d, phase_x S_x) (complex, re/im),
 * S_d = sum_L Cd_L I_dir(L), S_x = sum_L Cx_L I_ex(L) in ascending L, a term
 * with |I| < 1e-20 skipped (dwba_coupling.py:372-430).  For ionization these
 * are f_coeff and g_coeff.  n_combos <= 4096, L_max <= 63. */
int32_t dwba_amp_coeffs_f64(int32_t n_pairs, const dwba_amp_pair* d_pairs, int32_t n_combos, int32_t L_max,
                            const double* d_integrals, const double* d_coef, const double* d_phase, double* d_a,
                            void* stream);

/* F[slot][combo][j] = sum_p a_d[p][combo] Y_{l_f(p), m_dir[combo]}(theta_j),
 * G likewise with a_x and m_ex, over the pairs p in [slot_first[slot],
 * slot_first[slot+1]) in index order (complex, re/im interleaved).  m_dir,
 * m_ex: host arrays of n_combos values (M_i - M_f and M_f - M_i); d_ylm from
 * dwba_ylm_f64 with the same l_max, m_max, n_theta.  A pair whose l_f > l_max
 * contributes 0.  d_ylm NULL (with n_theta == 1): Y = 1, the coherent sums of
 * the coefficients over the slot's pairs (the ionization SDCS amplitudes,
 * ionization.py:631-642).  n_slots <= 65535, n_combos <= 64. */
int32_t dwba_amp_sum_f64(int32_t n_slots, const int32_t* d_slot_first, const dwba_amp_pair* d_pairs,
                         int32_t n_combos, const int32_t* h_m_dir, const int32_t* h_m_ex, int32_t l_max,
                         int32_t m_max, int32_t n_theta, const double* d_ylm, const double* d_a, double* d_F,
                         double* d_G, void* stream);

/* d_dcs[slot][j] = sum_combo max(0, prefac[slot] (|F-G|^2 3/4 + |F+G|^2 / 4) (2 pi)^4),
 * prefac = (k_f/k_i) N_eq/(2 L_i + 1) (dcs_dwba, sigma_total.py:147-177);
 * d_tcs[slot] = 2 pi trapezoid(sin(theta) dcs, theta)
 * (integrate_dcs_over_angles, sigma_total.py:179-217), atomic units. */
int32_t dwba_dcs_f64(int32_t n_slots, int32_t n_combos, int32_t n_theta, const double* d_theta, const double* d_F,
                     const double* d_G, const double* d_prefac, double* d_dcs, double* d_tcs, void* stream);

/* d_tcs[k] = 2 pi trapezoid(sin(theta) d_dcs[k][:], theta) for k < n_rows: the
 * integral alone, for a DCS formed elsewhere (integrate_dcs_over_angles,
 * sigma_total.py:179-217). */
int32_t dwba_sphere_trapz_f64(int32_t n_rows, int32_t n_theta, const double* d_theta, const double* d_dcs,
                              double* d_tcs, void* stream);

/* ---- partial-wave sums --------------------------------------------------
 * The projectile partial-wave loops around the integrals, batched: the
 * excitation loop over l_i with its post-add convergence snapshots
 * (driver.py:877-1152) and the ionization SDCS loop with its coherent sums
 * over l_i (ionization.py:526-705).  The convergence decisions themselves are
 * a few comparisons per l_i, made on the host from these outputs. */

/* Running sums over the slots of each energy: slots
 * [energy_first[e], energy_first[e+1]) are the energy's l_i in order, with
 * F, G laid out as dwba_amp_sum_f64 writes them; cumF/cumG[slot] = the sum
 * over the energy's slots up to and including it (the reference's
 * total_amplitudes after adding that l_i).  flags[slot] (caller-zeroed)
 * gets bit 0 when the slot's own F or G holds a non-finite value and bit 1
 * when the running total does (the loop's numerical-safety checks).
 * n_combos <= 1024. */
int32_t dwba_pw_cumulate_f64(int32_t n_energies, const int32_t* d_energy_first, int32_t n_combos, int32_t n_theta,
                             const double* d_F, const double* d_G, double* d_cumF, double* d_cumG, int32_t* d_flags,
                             void* stream);

/* Ionization spin sums: a run is the pairs (indices into d_a, whose layout
 * is dwba_amp_coeffs_f64's) of one (point, l_ej, l_f) in ascending l_i,
 * listed at d_run_pairs[run_first[run] .. run_first[run+1]).  Per pair and
 * projection pair, with spin(f, g) = |f+g|^2/4 + 3|f-g|^2/4:
 *   dcoh = spin(running sum over the run after it) - spin(before it)
 *   inc  = spin(its own coefficients)
 * (the incremental coherent sum and the incoherent per-l_i sum of
 * ionization.py:627-642); d_dcoh, d_inc: [pair][combo]. */
int32_t dwba_ion_spin_terms_f64(int32_t n_runs, const int32_t* d_run_first, const int32_t* d_run_pairs,
                                int32_t n_combos, const double* d_a, double* d_dcoh, double* d_inc, void* stream);

/* out1[slot] = sum over the pairs d_slot_pairs[slot_first[slot] ..
 * slot_first[slot+1]) and all projection pairs of x1[pair][combo] (same for
 * x2/out2), in a fixed order: the per-l_i totals of the ionization loop. */
int32_t dwba_slot_sums_f64(int32_t n_slots, const int32_t* d_slot_first, const int32_t* d_slot_pairs,
                           int32_t n_combos, const double* d_x1, const double* d_x2, double* d_out1, double* d_out2,
                           void* stream);

/* ---- continuum waves ----------------------------------------------------
 * Distorted waves chi_l(k, r) for many partial waves at once: the
 * deterministic paths of solve_continuum_wave (continuum.py:1439-2200),
 * which the drivers call once per (energy, channel, l) before the integrals
 * (precompute_continuum_waves, driver.py:486-540).  The host chooses, per
 * wave, what the reference chooses from r, U, k and l alone (analytic bypass,
 * first propagated index, match point, start values); these entry points do
 * every grid-sized step.  Arrays are [wave][ld], ld >= every wave's n. */
#define DWBA_CW_JOHNSON 0
#define DWBA_CW_NUMEROV 1
#define DWBA_CW_MAX_EVENTS 512

typedef struct dwba_cw_wave {
    const double* r;       /* grid, n points */
    const double* U;       /* channel potential on the grid */
    double k;              /* wave number (a.u.) */
    double r1_ratio;       /* Johnson: R_1 = psi_0 / psi_1 of the propagated range (clipped to +-1e10) */
    double chi0, chi1;     /* Numerov: the first two values */
    int32_t n, l;
    int32_t start;         /* first propagated index; chi = 0 below it */
    int32_t method;        /* DWBA_CW_JOHNSON or DWBA_CW_NUMEROV */
    int32_t renorm;        /* Numerov: rescale check every `renorm` steps */
    int32_t pad_;
} dwba_cw_wave;

/* Propagation: Johnson's ratio recursion with the backward product and the
 * division by the maximum (_johnson_log_derivative_solve,
 * continuum.py:1180-1332), or local-step Numerov (_numerov_propagate,
 * :985-1069), each with its periodic rescaling reproduced exactly; then
 * chi' by the 5-point Fornberg stencil on the propagated range
 * (_derivative_5point, :924-983).  d_status[w] = 1 if a wave rescaled more
 * than DWBA_CW_MAX_EVENTS times (its values are then not valid).
 * workspace >= dwba_cw_workspace_bytes(n_waves, ld). */
size_t dwba_cw_workspace_bytes(int32_t n_waves, int32_t ld);
int32_t dwba_cw_propagate_f64(int32_t n_waves, const dwba_cw_wave* d_waves, int32_t ld, double* d_chi, double* d_dchi,
                              void* d_work, size_t work_bytes, int32_t* d_status, void* stream);

typedef struct dwba_cw_finish {
    const double* r;
    const double* w_trapz;
    double k, delta;       /* phase shift of the asymptotic form */
    double eta, sigma;     /* Coulomb parameter and phase (ionic waves) */
    int32_t n, l;
    int32_t idx_match;     /* match point */
    int32_t tail;          /* first index of the least-squares tail */
    int32_t coulomb;       /* 1: Coulomb asymptotic form (|z_ion| >= 1e-3) */
    int32_t pad_;
} dwba_cw_finish;

/* d_sums[w][0..5] = sums over j in [tail, n) of s^2, s c, c^2, chi s, chi c
 * and max |chi|, with (s, c) = (sin, cos) of k r - l pi/2 (neutral) or of the
 * Coulomb phase rho + eta ln 2 rho - l pi/2 + sigma: the normal equations of
 * the tail fit chi ~ a sin + b cos (_fit_asymptotic_phase_neutral /
 * _coulomb, continuum.py:656-767). */
int32_t dwba_cw_tail_sums_f64(int32_t n_waves, const dwba_cw_finish* d_fin, int32_t ld, const double* d_chi,
                              double* d_sums, void* stream);

/* The stitched wave (continuum.py:2030-2126): d_out[w][j] = s chi[j] for
 * j <= idx_match, the asymptotic wave sqrt(2/pi) [j^_l cos d - n^_l sin d]
 * (or sin(th - d) for ions) beyond, s from chi, chi' and the asymptotic
 * wave's value and 5-point derivative at the match point; divided by its
 * RMS amplitude when that is off 1 by more than 5%.  d_scratch: [wave][ld]. */
int32_t dwba_cw_stitch_f64(int32_t n_waves, const dwba_cw_finish* d_fin, int32_t ld, const double* d_chi,
                           const double* d_dchi, double* d_scratch, double* d_out, void* stream);

typedef struct dwba_cw_bypass {
    const double* r;
    const double* w_trapz;
    const double* U;
    double k;
    int32_t n, l;
} dwba_cw_bypass;

/* The analytic bypass of a weak neutral channel (continuum.py:1620-1668):
 * d_out[w][j] = sqrt(2/pi) r_j j_l(k r_j) k and d_delta[w] = the Born phase
 * -k sum_j w_j U_j j_l(k r_j)^2 r_j^2 clipped to [-pi/2, pi/2]. */
int32_t dwba_cw_bypass_f64(int32_t n_waves, const dwba_cw_bypass* d_byp, int32_t ld, double* d_out, double* d_delta,
                           void* stream);

/* d_dst[k*dst_ld + i] = d_w[i] * d_src[k*src_ld + i] for i < n, 0 for
 * n <= i < dst_ld, k < n_rows (the quadrature-weighted densities
 * alpha = w rho_a, beta = w rho_b of the plain branch,
 * dwba_matrix_elements.py:1243-1245). */
int32_t dwba_weight_rows(const double* d_w, int32_t n, int32_t n_rows, const double* d_src, int32_t src_ld,
                         double* d_dst, int32_t dst_ld, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* DWBA_HIP_H */
