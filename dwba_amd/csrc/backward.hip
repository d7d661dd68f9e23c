// Gradient of the plain multipole contraction with respect to the densities.
//
// For pairs p = 0..P-1 on one grid (r, w) the forward quantity is
//     I_L(p) = sum_ij w_i rho_a[p][i] K_L(i, j) w_j rho_b[p][j]
// (the plain branch of radial_ME_all_L, dwba_matrix_elements.py:1243-1245,
// :1288-1289; the reference has no gradient).  For a loss sum_{p,L} g[p][L] I_L(p)
//     d/d rho_a[p][i] = w_i sum_j Keff_p(i, j) w_j rho_b[p][j]
//     d/d rho_b[p][i] = w_i sum_j Keff_p(i, j) w_j rho_a[p][j]      (K symmetric)
// with Keff_p = sum_L g[p][L] K_L.  Keff is regenerated element by element,
// never stored: with q = min(r_< / r_>, 1 - 1e-12) (q = 1 - 1e-12 on the
// diagonal, the reference's clamp) and K_0 = 1 / r_>,
//     Keff_p(i, j) = K_0 * (g_0 + q (g_1 + q (g_2 + ...)))        (Horner)
// so an element costs L_max + 1 FMAs plus two for the two gradients.
//
// Work mapping: workgroup = (256 rows, one pair); thread = row i; columns are
// staged through LDS 256 at a time and read as broadcasts.  The g[p][.] of
// the workgroup are uniform and live in scalar registers.  Type T is the
// arithmetic type of Keff and of the products: double, or float with each
// 256-column chunk summed in FP32 and the chunks added in FP64.
#include "common.hpp"

namespace dwba {
namespace {

constexpr int BT = 256;     // rows per workgroup
constexpr int BCH = 256;    // columns per LDS round

template <typename T> __device__ inline T kclamp();
template <> __device__ inline double kclamp<double>() { return 1.0 - 1e-12; }
template <> __device__ inline float kclamp<float>() { return 1.0f; }  // 1 - 1e-12 rounds to 1 in FP32

template <typename T> __device__ inline T tmin(T a, T b);
template <> __device__ inline double tmin<double>(double a, double b) { return fmin(a, b); }
template <> __device__ inline float tmin<float>(float a, float b) { return fminf(a, b); }

template <int NL, typename T>
__global__ __launch_bounds__(BT) void il_backward_kernel(const double* __restrict__ r, const double* __restrict__ w,
                                                         int n, int ld, const double* __restrict__ rho_a,
                                                         const double* __restrict__ rho_b,
                                                         const double* __restrict__ g, double* __restrict__ grad_a,
                                                         double* __restrict__ grad_b) {
    __shared__ T s_r[BCH], s_ir[BCH], s_a[BCH], s_b[BCH];
    const int p = blockIdx.y;
    const int i = blockIdx.x * BT + threadIdx.x;
    const bool row = i < n;
    const double ri_d = row ? r[i] : 1.0;
    const T ri = (T)ri_d, iri = (T)(1.0 / ri_d);
    const double* gp = g + (size_t)p * NL;
    T gc[NL];
#pragma unroll
    for (int L = 0; L < NL; ++L) gc[L] = (T)gp[L];
    const double* pa = rho_a + (size_t)p * ld;
    const double* pb = rho_b + (size_t)p * ld;
    const T one_m = kclamp<T>();
    double acc_a = 0.0, acc_b = 0.0;
    for (int j0 = 0; j0 < n; j0 += BCH) {
        __syncthreads();
        {
            const int j = j0 + threadIdx.x;
            const bool in = j < n;
            const double rj = in ? r[j] : 1.0;
            const double wj = in ? w[j] : 0.0;
            s_r[threadIdx.x] = (T)rj;
            s_ir[threadIdx.x] = (T)(1.0 / rj);
            s_a[threadIdx.x] = (T)(in ? wj * pa[j] : 0.0);
            s_b[threadIdx.x] = (T)(in ? wj * pb[j] : 0.0);
        }
        __syncthreads();
        const int jn = min(BCH, n - j0);
        T sa = (T)0, sb = (T)0;
#pragma unroll 2
        for (int c = 0; c < jn; ++c) {
            const int j = j0 + c;
            T q, k0;
            if (j > i) { k0 = s_ir[c]; q = tmin<T>(ri * s_ir[c], one_m); }
            else if (j < i) { k0 = iri; q = tmin<T>(s_r[c] * iri, one_m); }
            else { k0 = iri; q = one_m; }
            T h = gc[NL - 1];
#pragma unroll
            for (int L = NL - 2; L >= 0; --L) h = fma(h, q, gc[L]);
            const T k = h * k0;
            sa = fma(k, s_b[c], sa);
            sb = fma(k, s_a[c], sb);
        }
        acc_a += (double)sa;
        acc_b += (double)sb;
    }
    if (i < ld) {
        const double wi = row ? w[i] : 0.0;
        if (grad_a) grad_a[(size_t)p * ld + i] = row ? wi * acc_a : 0.0;
        if (grad_b) grad_b[(size_t)p * ld + i] = row ? wi * acc_b : 0.0;
    }
}

// dst[k][i] = w[i] * src[k][i] for i < n, 0 for n <= i < dst_ld.
__global__ __launch_bounds__(256) void weight_rows_kernel(const double* __restrict__ w, int n,
                                                          const double* __restrict__ src, int src_ld,
                                                          double* __restrict__ dst, int dst_ld, int64_t total) {
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (k >= total) return;
    const int64_t row = k / dst_ld;
    const int i = (int)(k - row * dst_ld);
    dst[k] = i < n ? w[i] * src[row * src_ld + i] : 0.0;
}

template <typename T>
int launch_backward(const char* name, const double* d_r, const double* d_w, int32_t n, int32_t n_pairs, int32_t ld,
                    const double* d_rho_a, const double* d_rho_b, const double* d_g, int32_t L_max, double* d_grad_a,
                    double* d_grad_b, void* stream) {
    DWBA_REQUIRE(initialised(), "%s: call dwba_init first", name);
    DWBA_REQUIRE(d_r && d_w && d_rho_a && d_rho_b && d_g, "%s: null input pointer", name);
    DWBA_REQUIRE(d_grad_a || d_grad_b, "%s: both outputs are NULL", name);
    DWBA_REQUIRE(n > 0 && ld >= n && n_pairs > 0 && n_pairs <= 65535, "%s: bad sizes n=%d ld=%d n_pairs=%d", name, n,
                 ld, n_pairs);
    DWBA_REQUIRE(L_max >= 0 && L_max <= 63, "%s: L_max %d outside [0, 63]", name, L_max);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    dim3 grid((ld + BT - 1) / BT, n_pairs);
    switch (L_max + 1) {
#define C(NLV) case NLV: hipLaunchKernelGGL((il_backward_kernel<NLV, T>), grid, dim3(BT), 0, s, d_r, d_w, n, ld, \
                                           d_rho_a, d_rho_b, d_g, d_grad_a, d_grad_b); break;
        C(1) C(2) C(3) C(4) C(5) C(6) C(7) C(8) C(9) C(10) C(11) C(12) C(13) C(14) C(15) C(16)
        C(17) C(18) C(19) C(20) C(21) C(22) C(23) C(24) C(25) C(26) C(27) C(28) C(29) C(30) C(31) C(32)
        C(33) C(34) C(35) C(36) C(37) C(38) C(39) C(40) C(41) C(42) C(43) C(44) C(45) C(46) C(47) C(48)
        C(49) C(50) C(51) C(52) C(53) C(54) C(55) C(56) C(57) C(58) C(59) C(60) C(61) C(62) C(63) C(64)
#undef C
        default: return set_error(DWBA_ERR_ARG, "%s: L_max %d", name, L_max);
    }
    DWBA_LAUNCH_CHECK(name);
    return DWBA_OK;
}

}  // namespace
}  // namespace dwba

using namespace dwba;

extern "C" {

int32_t dwba_il_backward_f64(const double* d_r, const double* d_w, int32_t n, int32_t n_pairs, int32_t ld,
                             const double* d_rho_a, const double* d_rho_b, const double* d_g, int32_t L_max,
                             double* d_grad_a, double* d_grad_b, void* stream) {
    return launch_backward<double>("dwba_il_backward_f64", d_r, d_w, n, n_pairs, ld, d_rho_a, d_rho_b, d_g, L_max,
                                   d_grad_a, d_grad_b, stream);
}

int32_t dwba_il_backward_f32(const double* d_r, const double* d_w, int32_t n, int32_t n_pairs, int32_t ld,
                             const double* d_rho_a, const double* d_rho_b, const double* d_g, int32_t L_max,
                             double* d_grad_a, double* d_grad_b, void* stream) {
    return launch_backward<float>("dwba_il_backward_f32", d_r, d_w, n, n_pairs, ld, d_rho_a, d_rho_b, d_g, L_max,
                                  d_grad_a, d_grad_b, stream);
}

int32_t dwba_weight_rows(const double* d_w, int32_t n, int32_t n_rows, const double* d_src, int32_t src_ld,
                         double* d_dst, int32_t dst_ld, void* stream) {
    DWBA_REQUIRE(initialised(), "dwba_weight_rows: call dwba_init first");
    DWBA_REQUIRE(d_w && d_src && d_dst && n > 0 && src_ld >= n && dst_ld >= n && n_rows > 0,
                 "dwba_weight_rows: bad arguments");
    const int64_t total = (int64_t)n_rows * dst_ld;
    hipLaunchKernelGGL(weight_rows_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), d_w, n, d_src, src_ld, d_dst, dst_ld, total);
    DWBA_LAUNCH_CHECK("weight_rows_kernel");
    return DWBA_OK;
}

}  // extern "C"
