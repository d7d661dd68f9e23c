/*
 * dwba_hip.h -- C ABI of the MI355X (gfx950) radial multipole integral library.
 *
 * I_L = sum_ij alpha_i K_L(r_i, r_j) beta_j,  K_L = r_<^L / r_>^(L+1)
 * (diagonal (1-1e-12)^L / r_i), for L = 0..L_max, direct and exchange.
 *
 * Conventions
 *  - Every pointer argument named d_* is DEVICE memory owned by the caller;
 *    h_* is host memory.  The library allocates nothing behind the caller's
 *    back except the per-process state created by dwba_init.
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

#ifdef __cplusplus
}
#endif
#endif /* DWBA_HIP_H */
