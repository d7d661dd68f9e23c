// Amplitude assembly from the radial integrals, FP64, on the device.
//
// Replaces the per-(M_i, M_f) host loops that consume the I_L dicts:
// calculate_amplitude_contribution (dwba_coupling.py:285-504) as the driver
// accumulates it over the channel pairs of an energy (driver.py:1041-1052),
// calculate_ionization_coefficients (dwba_coupling.py:507-634), dcs_dwba and
// integrate_dcs_over_angles (sigma_total.py:147-217).  The angular
// coefficients Cd_L, Cx_L and the constant phases of each channel type come
// from the host (exact Racah sums); everything that touches the integrals or
// the angle grid runs here:
//
//   ylm_kernel         Y_{l,m}(theta_j, 0), Condon-Shortley phase, by the
//                      fully normalised three-term recurrence in l
//   amp_coeff_kernel   a_dir = scale phase_d sum_L Cd_L I_dir(L),
//                      a_ex = scale phase_x sum_L Cx_L I_ex(L)   per (pair, projection pair)
//   amp_sum_kernel     F(theta) = sum_pairs a_dir Y_{l_f, m_dir}(theta),
//                      G(theta) = sum_pairs a_ex Y_{l_f, m_ex}(theta)   per (energy, projection pair)
//   dcs_kernel         DCS = sum_proj (k_f/k_i) N_eq/(2L_i+1) [|F+G|^2/4 + 3|F-G|^2/4] (2 pi)^4,
//                      TCS = 2 pi trapezoid(sin theta DCS)
#include "common.hpp"

namespace dwba {
namespace {

constexpr int AT = 256;
constexpr double PI_A = 3.14159265358979323846;
constexpr int MAX_COMBOS = 64;

struct ComboMs {
    int m_dir[MAX_COMBOS];   // m of Y_{l_f, m} in F (= M_i - M_f)
    int m_ex[MAX_COMBOS];    // m of Y_{l_f, m} in G (= M_f - M_i)
};

// Y_{l,m}(theta, 0) for l <= l_max and m = -m_max..m_max (0 where |m| > l):
// ylm[(l * (2 m_max + 1) + m + m_max) * n_theta + j].  One thread per (theta_j, m >= 0);
// Y_{l,-m} = (-1)^m Y_{l,m} at phi = 0.
__global__ __launch_bounds__(AT) void ylm_kernel(int l_max, int m_max, int n_theta, const double* __restrict__ theta,
                                                 double* __restrict__ ylm) {
    const int j = blockIdx.x * AT + threadIdx.x;
    const int m = blockIdx.y;
    if (j >= n_theta) return;
    const int nm = 2 * m_max + 1;
    auto put = [&](int l, double v) {
        ylm[((size_t)l * nm + (m + m_max)) * n_theta + j] = v;
        if (m > 0) ylm[((size_t)l * nm + (m_max - m)) * n_theta + j] = (m & 1) ? -v : v;
    };
    for (int l = 0; l < m && l <= l_max; ++l) put(l, 0.0);
    if (m > l_max) return;
    double s, c;
    sincos(theta[j], &s, &c);
    double ymm = 1.0 / sqrt(4.0 * PI_A);                       // Y_00
    for (int k = 1; k <= m; ++k) ymm *= -sqrt((2.0 * k + 1.0) / (2.0 * k)) * s;
    put(m, ymm);
    if (m + 1 > l_max) return;
    double y0 = ymm, y1 = sqrt(2.0 * m + 3.0) * c * ymm;       // Y_{m+1, m}
    put(m + 1, y1);
    for (int l = m + 2; l <= l_max; ++l) {
        const double ll = (double)l, mm = (double)m;
        const double a = sqrt((4.0 * ll * ll - 1.0) / (ll * ll - mm * mm));
        const double b = sqrt(((ll - 1.0) * (ll - 1.0) - mm * mm) / (4.0 * (ll - 1.0) * (ll - 1.0) - 1.0));
        const double y2 = a * (c * y1 - b * y0);
        put(l, y2);
        y0 = y1;
        y1 = y2;
    }
}

// One thread per (pair, projection pair): the L sums in ascending L, a term
// with |I_L| < 1e-20 or a zero coefficient skipped (the reference's cuts).
__global__ __launch_bounds__(AT) void amp_coeff_kernel(int n_pairs, const dwba_amp_pair* __restrict__ pairs,
                                                       int n_combos, int L_max, const double* __restrict__ integrals,
                                                       const double* __restrict__ coef,
                                                       const double* __restrict__ phase, double* __restrict__ a) {
    const long x = (long)blockIdx.x * AT + threadIdx.x;
    if (x >= (long)n_pairs * n_combos) return;
    const int p = (int)(x / n_combos), c = (int)(x - (long)p * n_combos);
    const dwba_amp_pair P = pairs[p];
    const int nl = L_max + 1;
    const double* __restrict__ I = integrals + (size_t)P.row * 2 * nl;
    const double* __restrict__ C = coef + ((size_t)P.type * n_combos + c) * 2 * nl;
    const double* __restrict__ ph = phase + ((size_t)P.type * n_combos + c) * 4;
    double sd = 0.0, sx = 0.0;
    for (int L = 0; L < nl; ++L) {
        const double Id = I[L], Ix = I[nl + L];
        const double cd = C[L], cx = C[nl + L];
        if (cd != 0.0 && !(fabs(Id) < 1e-20)) sd += cd * Id;
        if (cx != 0.0 && !(fabs(Ix) < 1e-20)) sx += cx * Ix;
    }
    const double vd = P.scale * sd, vx = P.scale * sx;
    double* __restrict__ o = a + (size_t)x * 4;
    o[0] = vd * ph[0];
    o[1] = vd * ph[1];
    o[2] = vx * ph[2];
    o[3] = vx * ph[3];
}

// F, G of one (energy, projection pair) over a block of angles: pairs of the
// energy in their given order, one angle per thread.
__global__ __launch_bounds__(AT) void amp_sum_kernel(const int* __restrict__ slot_first,
                                                     const dwba_amp_pair* __restrict__ pairs, int n_combos,
                                                     ComboMs ms, int l_max, int m_max, int n_theta,
                                                     const double* __restrict__ ylm, const double* __restrict__ a,
                                                     double* __restrict__ F, double* __restrict__ G) {
    const int j = blockIdx.x * AT + threadIdx.x;
    const int slot = blockIdx.y, c = blockIdx.z;
    if (j >= n_theta) return;
    const int nm = 2 * m_max + 1;
    const int md = ms.m_dir[c], mx = ms.m_ex[c];
    const bool okd = md >= -m_max && md <= m_max, okx = mx >= -m_max && mx <= m_max;
    double fr = 0.0, fi = 0.0, gr = 0.0, gi = 0.0;
    const int p1 = slot_first[slot + 1];
    for (int p = slot_first[slot]; p < p1; ++p) {
        const int lf = pairs[p].l_f;
        const double* __restrict__ ap = a + ((size_t)p * n_combos + c) * 4;
        double yd = 1.0, yx = 1.0;                 // no table: the coherent sums of the coefficients themselves
        if (ylm) {
            const bool inl = lf >= 0 && lf <= l_max;
            yd = (okd && inl) ? ylm[((size_t)lf * nm + md + m_max) * n_theta + j] : 0.0;
            yx = (okx && inl) ? ylm[((size_t)lf * nm + mx + m_max) * n_theta + j] : 0.0;
        }
        fr = fma(ap[0], yd, fr);
        fi = fma(ap[1], yd, fi);
        gr = fma(ap[2], yx, gr);
        gi = fma(ap[3], yx, gi);
    }
    const size_t o = (((size_t)slot * n_combos + c) * n_theta + j) * 2;
    F[o] = fr;
    F[o + 1] = fi;
    G[o] = gr;
    G[o + 1] = gi;
}

// 2 pi sum over the workgroup's partial trapezoid sums (thread 0 writes).
__device__ inline void block_sphere_sum(double acc, double* s_red, double* out) {
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < AT / 64; ++w) s += s_red[w];
        *out = 2.0 * PI_A * s;
    }
}

// 2 pi trapezoid(sin theta dcs) of given DCS rows: one workgroup per row.
__global__ __launch_bounds__(AT) void sphere_trapz_kernel(int n_theta, const double* __restrict__ theta,
                                                          const double* __restrict__ dcs, double* __restrict__ tcs) {
    __shared__ double s_red[AT / 64];
    const double* __restrict__ d = dcs + (size_t)blockIdx.x * n_theta;
    double acc = 0.0;
    for (int j = threadIdx.x; j + 1 < n_theta; j += AT)
        acc += 0.5 * (sin(theta[j]) * d[j] + sin(theta[j + 1]) * d[j + 1]) * (theta[j + 1] - theta[j]);
    block_sphere_sum(acc, s_red, tcs + blockIdx.x);
}

// DCS of one energy per angle and its integral over the sphere: one workgroup per energy.
__global__ __launch_bounds__(AT) void dcs_kernel(int n_combos, int n_theta, const double* __restrict__ theta,
                                                 const double* __restrict__ F, const double* __restrict__ G,
                                                 const double* __restrict__ prefac, double* __restrict__ dcs,
                                                 double* __restrict__ tcs) {
    __shared__ double s_red[AT / 64];
    const int slot = blockIdx.x;
    const double pf = prefac[slot];
    const double two_pi4 = (2.0 * PI_A) * (2.0 * PI_A) * (2.0 * PI_A) * (2.0 * PI_A);
    auto d_at = [&](int j) {
        double d = 0.0;
        for (int c = 0; c < n_combos; ++c) {
            const size_t o = (((size_t)slot * n_combos + c) * n_theta + j) * 2;
            const double fr = F[o], fi = F[o + 1], gr = G[o], gi = G[o + 1];
            const double pr = fr + gr, pi = fi + gi, mr = fr - gr, mi = fi - gi;
            const double v = pf * (0.75 * (mr * mr + mi * mi) + 0.25 * (pr * pr + pi * pi)) * two_pi4;
            d += v > 0.0 ? v : 0.0;
        }
        return d;
    };
    double acc = 0.0;
    for (int j = threadIdx.x; j < n_theta; j += AT) {
        const double dj = d_at(j);
        dcs[(size_t)slot * n_theta + j] = dj;
        if (j + 1 < n_theta) {
            const double d1 = d_at(j + 1);
            acc += 0.5 * (sin(theta[j]) * dj + sin(theta[j + 1]) * d1) * (theta[j + 1] - theta[j]);
        }
    }
    block_sphere_sum(acc, s_red, tcs + slot);
}

// Partial-wave sums of an energy: slot k of an energy holds the amplitudes of
// one projectile wave l_i (its F, G); the running totals after each slot are
// the snapshots the partial-wave loop tests for convergence.  One thread per
// (energy, projection pair, angle) runs over the energy's slots in order;
// flags[slot] |= 1 when the slot's own F or G has a non-finite value, 2 when
// the running total after it has.
__global__ __launch_bounds__(AT) void pw_cumulate_kernel(const int* __restrict__ energy_first, int n_combos,
                                                         int n_theta, const double* __restrict__ F,
                                                         const double* __restrict__ G, double* __restrict__ cF,
                                                         double* __restrict__ cG, int* __restrict__ flags) {
    const int j = blockIdx.x * AT + threadIdx.x;
    const int e = blockIdx.y, c = blockIdx.z;
    if (j >= n_theta) return;
    double fr = 0.0, fi = 0.0, gr = 0.0, gi = 0.0;
    const int s1 = energy_first[e + 1];
    for (int s = energy_first[e]; s < s1; ++s) {
        const size_t o = (((size_t)s * n_combos + c) * n_theta + j) * 2;
        const double a = F[o], b = F[o + 1], x = G[o], y = G[o + 1];
        fr += a;
        fi += b;
        gr += x;
        gi += y;
        cF[o] = fr;
        cF[o + 1] = fi;
        cG[o] = gr;
        cG[o + 1] = gi;
        int fl = 0;
        if (!isfinite(a) || !isfinite(b) || !isfinite(x) || !isfinite(y)) fl |= 1;
        if (!isfinite(fr) || !isfinite(fi) || !isfinite(gr) || !isfinite(gi)) fl |= 2;
        if (fl) atomicOr(flags + s, fl);
    }
}

__device__ inline double spin_combo(double fr, double fi, double gr, double gi) {
    const double pr = fr + gr, pi = fi + gi, mr = fr - gr, mi = fi - gi;
    return 0.25 * (pr * pr + pi * pi) + 0.75 * (mr * mr + mi * mi);
}

// Ionization partial waves.  A run is the pairs of one (point, l_ej, l_f) in
// ascending l_i; per projection pair its coefficients (f, g) accumulate over
// the run (the coherent SDCS amplitude of that key).  For every pair:
//   dcoh = spin(running sum after it) - spin(running sum before it),
//   inc  = spin(its own coefficients),       spin(f, g) = |f+g|^2/4 + 3|f-g|^2/4.
// One thread per (run, projection pair).
__global__ __launch_bounds__(AT) void ion_spin_terms_kernel(int n_runs, const int* __restrict__ run_first,
                                                            const int* __restrict__ run_pairs, int n_combos,
                                                            const double* __restrict__ a, double* __restrict__ dcoh,
                                                            double* __restrict__ inc) {
    const long x = (long)blockIdx.x * AT + threadIdx.x;
    if (x >= (long)n_runs * n_combos) return;
    const int run = (int)(x / n_combos), c = (int)(x - (long)run * n_combos);
    double fr = 0.0, fi = 0.0, gr = 0.0, gi = 0.0, prev = 0.0;
    for (int q = run_first[run]; q < run_first[run + 1]; ++q) {
        const size_t o = (size_t)run_pairs[q] * n_combos + c;
        const double* __restrict__ ap = a + o * 4;
        fr += ap[0];
        fi += ap[1];
        gr += ap[2];
        gi += ap[3];
        const double now = spin_combo(fr, fi, gr, gi);
        dcoh[o] = now - prev;
        inc[o] = spin_combo(ap[0], ap[1], ap[2], ap[3]);
        prev = now;
    }
}

// out1[slot] = sum over the slot's pairs (in the listed order) and projection
// pairs of x1[pair][combo]; the same for x2.  One wavefront per slot.
__global__ __launch_bounds__(64) void slot_sums_kernel(const int* __restrict__ slot_first,
                                                       const int* __restrict__ slot_pairs, int n_combos,
                                                       const double* __restrict__ x1, const double* __restrict__ x2,
                                                       double* __restrict__ out1, double* __restrict__ out2) {
    const int slot = blockIdx.x, lane = threadIdx.x;
    const int q0 = slot_first[slot], q1 = slot_first[slot + 1];
    double s1 = 0.0, s2 = 0.0;
    for (int q = q0; q < q1; ++q) {
        const size_t base = (size_t)slot_pairs[q] * n_combos;
        for (int c = lane; c < n_combos; c += 64) {
            s1 += x1[base + c];
            s2 += x2[base + c];
        }
    }
    s1 = wave_sum(s1);
    s2 = wave_sum(s2);
    if (lane == 0) {
        out1[slot] = s1;
        out2[slot] = s2;
    }
}

}  // namespace
}  // namespace dwba

using namespace dwba;

extern "C" {

int32_t dwba_ylm_f64(int32_t l_max, int32_t m_max, int32_t n_theta, const double* d_theta, double* d_ylm,
                     void* stream) {
    DWBA_REQUIRE(initialised(), "dwba_ylm_f64: call dwba_init first");
    DWBA_REQUIRE(l_max >= 0 && l_max <= 1000 && m_max >= 0 && m_max <= 1000 && n_theta >= 1 && d_theta && d_ylm,
                 "dwba_ylm_f64: bad arguments (l_max=%d, m_max=%d, n_theta=%d)", l_max, m_max, n_theta);
    hipLaunchKernelGGL(ylm_kernel, dim3((n_theta + AT - 1) / AT, m_max + 1), dim3(AT), 0,
                       reinterpret_cast<hipStream_t>(stream), l_max, m_max, n_theta, d_theta, d_ylm);
    DWBA_LAUNCH_CHECK("ylm_kernel");
    return DWBA_OK;
}

int32_t dwba_amp_coeffs_f64(int32_t n_pairs, const dwba_amp_pair* d_pairs, int32_t n_combos, int32_t L_max,
                            const double* d_integrals, const double* d_coef, const double* d_phase, double* d_a,
                            void* stream) {
    DWBA_REQUIRE(initialised(), "dwba_amp_coeffs_f64: call dwba_init first");
    DWBA_REQUIRE(n_pairs >= 1 && n_combos >= 1 && n_combos <= 4096 && L_max >= 0 && L_max <= 63 && d_pairs &&
                 d_integrals && d_coef && d_phase && d_a, "dwba_amp_coeffs_f64: bad arguments");
    const long n = (long)n_pairs * n_combos;
    hipLaunchKernelGGL(amp_coeff_kernel, dim3((unsigned)((n + AT - 1) / AT)), dim3(AT), 0,
                       reinterpret_cast<hipStream_t>(stream), n_pairs, d_pairs, n_combos, L_max, d_integrals, d_coef,
                       d_phase, d_a);
    DWBA_LAUNCH_CHECK("amp_coeff_kernel");
    return DWBA_OK;
}

int32_t dwba_amp_sum_f64(int32_t n_slots, const int32_t* d_slot_first, const dwba_amp_pair* d_pairs,
                         int32_t n_combos, const int32_t* h_m_dir, const int32_t* h_m_ex, int32_t l_max,
                         int32_t m_max, int32_t n_theta, const double* d_ylm, const double* d_a, double* d_F,
                         double* d_G, void* stream) {
    DWBA_REQUIRE(initialised(), "dwba_amp_sum_f64: call dwba_init first");
    DWBA_REQUIRE(n_slots >= 1 && n_slots <= 65535 && n_combos >= 1 && n_combos <= MAX_COMBOS && n_theta >= 1 &&
                 d_slot_first && d_pairs && h_m_dir && h_m_ex && (d_ylm || n_theta == 1) && d_a && d_F && d_G &&
                 l_max >= 0 && m_max >= 0, "dwba_amp_sum_f64: bad arguments");
    ComboMs ms{};
    for (int c = 0; c < n_combos; ++c) {
        ms.m_dir[c] = h_m_dir[c];
        ms.m_ex[c] = h_m_ex[c];
    }
    hipLaunchKernelGGL(amp_sum_kernel, dim3((n_theta + AT - 1) / AT, n_slots, n_combos), dim3(AT), 0,
                       reinterpret_cast<hipStream_t>(stream), d_slot_first, d_pairs, n_combos, ms, l_max, m_max,
                       n_theta, d_ylm, d_a, d_F, d_G);
    DWBA_LAUNCH_CHECK("amp_sum_kernel");
    return DWBA_OK;
}

int32_t dwba_dcs_f64(int32_t n_slots, int32_t n_combos, int32_t n_theta, const double* d_theta, const double* d_F,
                     const double* d_G, const double* d_prefac, double* d_dcs, double* d_tcs, void* stream) {
    DWBA_REQUIRE(initialised(), "dwba_dcs_f64: call dwba_init first");
    DWBA_REQUIRE(n_slots >= 1 && n_combos >= 1 && n_theta >= 1 && d_theta && d_F && d_G && d_prefac && d_dcs && d_tcs,
                 "dwba_dcs_f64: bad arguments");
    hipLaunchKernelGGL(dcs_kernel, dim3(n_slots), dim3(AT), 0, reinterpret_cast<hipStream_t>(stream), n_combos,
                       n_theta, d_theta, d_F, d_G, d_prefac, d_dcs, d_tcs);
    DWBA_LAUNCH_CHECK("dcs_kernel");
    return DWBA_OK;
}

int32_t dwba_pw_cumulate_f64(int32_t n_energies, const int32_t* d_energy_first, int32_t n_combos, int32_t n_theta,
                             const double* d_F, const double* d_G, double* d_cumF, double* d_cumG, int32_t* d_flags,
                             void* stream) {
    DWBA_REQUIRE(initialised(), "dwba_pw_cumulate_f64: call dwba_init first");
    DWBA_REQUIRE(n_energies >= 1 && n_energies <= 65535 && n_combos >= 1 && n_combos <= 1024 && n_theta >= 1 &&
                 d_energy_first && d_F && d_G && d_cumF && d_cumG && d_flags, "dwba_pw_cumulate_f64: bad arguments");
    hipLaunchKernelGGL(pw_cumulate_kernel, dim3((n_theta + AT - 1) / AT, n_energies, n_combos), dim3(AT), 0,
                       reinterpret_cast<hipStream_t>(stream), d_energy_first, n_combos, n_theta, d_F, d_G, d_cumF,
                       d_cumG, d_flags);
    DWBA_LAUNCH_CHECK("pw_cumulate_kernel");
    return DWBA_OK;
}

int32_t dwba_ion_spin_terms_f64(int32_t n_runs, const int32_t* d_run_first, const int32_t* d_run_pairs,
                                int32_t n_combos, const double* d_a, double* d_dcoh, double* d_inc, void* stream) {
    DWBA_REQUIRE(initialised(), "dwba_ion_spin_terms_f64: call dwba_init first");
    DWBA_REQUIRE(n_runs >= 1 && n_combos >= 1 && n_combos <= 4096 && d_run_first && d_run_pairs && d_a && d_dcoh &&
                 d_inc, "dwba_ion_spin_terms_f64: bad arguments");
    const long n = (long)n_runs * n_combos;
    hipLaunchKernelGGL(ion_spin_terms_kernel, dim3((unsigned)((n + AT - 1) / AT)), dim3(AT), 0,
                       reinterpret_cast<hipStream_t>(stream), n_runs, d_run_first, d_run_pairs, n_combos, d_a, d_dcoh,
                       d_inc);
    DWBA_LAUNCH_CHECK("ion_spin_terms_kernel");
    return DWBA_OK;
}

int32_t dwba_slot_sums_f64(int32_t n_slots, const int32_t* d_slot_first, const int32_t* d_slot_pairs,
                           int32_t n_combos, const double* d_x1, const double* d_x2, double* d_out1, double* d_out2,
                           void* stream) {
    DWBA_REQUIRE(initialised(), "dwba_slot_sums_f64: call dwba_init first");
    DWBA_REQUIRE(n_slots >= 1 && n_combos >= 1 && d_slot_first && d_slot_pairs && d_x1 && d_x2 && d_out1 && d_out2,
                 "dwba_slot_sums_f64: bad arguments");
    hipLaunchKernelGGL(slot_sums_kernel, dim3(n_slots), dim3(64), 0, reinterpret_cast<hipStream_t>(stream),
                       d_slot_first, d_slot_pairs, n_combos, d_x1, d_x2, d_out1, d_out2);
    DWBA_LAUNCH_CHECK("slot_sums_kernel");
    return DWBA_OK;
}

int32_t dwba_sphere_trapz_f64(int32_t n_rows, int32_t n_theta, const double* d_theta, const double* d_dcs,
                              double* d_tcs, void* stream) {
    DWBA_REQUIRE(initialised(), "dwba_sphere_trapz_f64: call dwba_init first");
    DWBA_REQUIRE(n_rows >= 1 && n_theta >= 1 && d_theta && d_dcs && d_tcs, "dwba_sphere_trapz_f64: bad arguments");
    hipLaunchKernelGGL(sphere_trapz_kernel, dim3(n_rows), dim3(AT), 0, reinterpret_cast<hipStream_t>(stream),
                       n_theta, d_theta, d_dcs, d_tcs);
    DWBA_LAUNCH_CHECK("sphere_trapz_kernel");
    return DWBA_OK;
}

}  // extern "C"
