// Continuum (distorted) waves chi_l(k, r) on the radial grid, batched over
// partial waves, FP64: the deterministic paths of solve_continuum_wave
// (continuum.py:1439-2200).
//
//   cw_propagate_kernel  one lane per wave over its whole range, serially:
//                        Johnson's ratio recursion R_{i+1} = 1/(T_i - R_i) and
//                        the backward product chi_{i-1} = R_i chi_i
//                        (continuum.py:1180-1332), or the local-step Numerov
//                        recursion (:985-1069).  Both rescale all values
//                        computed so far now and then; the lane records each
//                        rescale (position, factor) instead of touching the
//                        stored values, and
//   cw_apply_kernel      applies them afterwards, element by element, in the
//                        order the reference applies them (so the products
//                        round the same way), then Johnson's final division
//                        by the maximum;
//   cw_deriv_kernel      chi' by Fornberg's weights on the local 5 points (:924-983);
//   cw_tail_sums_kernel  the normal-equation sums of the tail least-squares
//                        phase fit (:656-767);
//   cw_stitch_kernel     the asymptotic wave A [j^_l cos d - n^_l sin d] (or its
//                        Coulomb form) beyond the match point, the scale from
//                        chi and chi' there, the stitched wave and the RMS
//                        amplitude check (:2030-2126);
//   cw_bypass_kernel     the analytic bypass: Born phase and Riccati-Bessel
//                        wave on the whole grid (:1620-1690).
// Spherical Bessel functions: upward recurrence where x > l, Miller's
// downward recurrence (normalised by j_0 or j_1) where x <= l; y_l upward.
#include "common.hpp"

#include <cmath>

namespace dwba {
namespace {

constexpr int CT = 256;
constexpr int CW_B = 16;         // grid points per register block of the serial recursions
constexpr double PI_C = 3.14159265358979323846;

struct CwWork {       // per-wave workspace (see dwba_cw_workspace_bytes)
    double* R;        // Johnson ratios / Numerov raw values: ld doubles
    double* raw;      // raw (unscaled) chi values: ld doubles
    int* ev_pos;      // rescale events
    double* ev_scale;
    int* ev_count;
    double* maxv;
};

// Per-wave workspace: R and raw values (2 ld doubles), rescale factors and the
// final maximum (DWBA_CW_MAX_EVENTS + 1 doubles), rescale positions and their
// count (DWBA_CW_MAX_EVENTS + 1 ints), rounded up to 256 bytes.
__host__ __device__ inline size_t cw_per_wave(int ld) {
    const size_t b = (size_t)(2 * ld + DWBA_CW_MAX_EVENTS + 1) * sizeof(double) +
                     (size_t)(DWBA_CW_MAX_EVENTS + 1) * sizeof(int);
    return (b + 255) / 256 * 256;
}

__device__ inline CwWork work_of(char* base, int w, int ld) {
    const size_t per = cw_per_wave(ld);
    char* p = base + (size_t)w * per;
    CwWork W;
    W.R = reinterpret_cast<double*>(p);
    W.raw = W.R + ld;
    W.ev_scale = W.raw + ld;
    W.maxv = W.ev_scale + DWBA_CW_MAX_EVENTS;
    W.ev_pos = reinterpret_cast<int*>(W.maxv + 1);
    W.ev_count = W.ev_pos + DWBA_CW_MAX_EVENTS;
    return W;
}

__device__ inline double q_of(double ell, double r, double U, double k2) {
    return __dsub_rn(__dadd_rn(__ddiv_rn(__dmul_rn(ell, ell + 1.0), __dmul_rn(r, r)), __dmul_rn(2.0, U)), k2);
}

// One lane per wave: the serial recursion over the wave's range [start, n).
__global__ __launch_bounds__(64) void cw_propagate_kernel(int n_waves, const dwba_cw_wave* __restrict__ waves, int ld,
                                                          char* __restrict__ work, int* __restrict__ status) {
    const int w = blockIdx.x * 64 + threadIdx.x;
    if (w >= n_waves) return;
    const dwba_cw_wave W = waves[w];
    CwWork S = work_of(work, w, ld);
    const int m = W.n - W.start;
    const double* __restrict__ r = W.r + W.start;
    const double* __restrict__ U = W.U + W.start;
    const double ell = (double)W.l, k2 = W.k * W.k;
    int nev = 0;
    bool over = false;
    if (W.method == DWBA_CW_JOHNSON) {
        double* __restrict__ R = S.R;
        R[0] = 0.0;
        R[1] = W.r1_ratio;
        double Ri = W.r1_ratio;
        double rm1 = r[0], r0 = r[1];
        // steps i = 1 .. m-2 in blocks of CW_B: the block's grid values are
        // loaded before its serial steps, so the loads are not on the chain
        for (int i0 = 1; i0 < m - 1; i0 += CW_B) {
            double rb[CW_B], ub[CW_B], Rb[CW_B];
#pragma unroll
            for (int b = 0; b < CW_B; ++b) {
                const int i = i0 + b;
                rb[b] = i + 1 < m ? r[i + 1] : 0.0;
                ub[b] = i < m ? U[i] : 0.0;
            }
#pragma unroll
            for (int b = 0; b < CW_B; ++b) {
                const int i = i0 + b;
                if (i < m - 1) {
                    const double rp1 = rb[b];
                    const double he = sqrt(__dmul_rn(__dsub_rn(rp1, r0), __dsub_rn(r0, rm1)));
                    const double Q = q_of(ell, r0, ub[b], k2);
                    double d = __dsub_rn(__dsub_rn(2.0, __dmul_rn(__dmul_rn(he, he), Q)), Ri);
                    if (fabs(d) < 1e-12) d = d != 0.0 ? copysign(1e-12, d) : 1e-12;
                    double Rn = __ddiv_rn(1.0, d);
                    Rn = fmin(fmax(Rn, -1e10), 1e10);
                    Rb[b] = Rn;
                    Ri = Rn;
                    rm1 = r0;
                    r0 = rp1;
                }
            }
#pragma unroll
            for (int b = 0; b < CW_B; ++b)
                if (i0 + b < m - 1) R[i0 + b + 1] = Rb[b];
        }
        // backward product; chi_{m-1} = 1, blocks of CW_B ratios loaded ahead
        double* __restrict__ raw = S.raw;
        raw[m - 1] = 1.0;
        double cur = 1.0, tmax = 1.0;
        for (int i0 = m - 1; i0 >= 1; i0 -= CW_B) {
            double Rb[CW_B], vb[CW_B];
#pragma unroll
            for (int b = 0; b < CW_B; ++b) Rb[b] = i0 - b >= 1 ? R[i0 - b] : 0.0;
#pragma unroll
            for (int b = 0; b < CW_B; ++b) {
                const int i = i0 - b;
                if (i >= 1) {
                    const double v = __dmul_rn(Rb[b], cur);
                    vb[b] = v;
                    cur = v;
                    tmax = fmax(tmax, fabs(v));
                    if (i % 100 == 0 && (tmax > 1e30 || (tmax < 1e-30 && tmax > 0.0))) {
                        const double s = __ddiv_rn(1.0, tmax);
                        if (nev < DWBA_CW_MAX_EVENTS) {
                            S.ev_pos[nev] = i - 1;  // applies to positions >= i-1
                            S.ev_scale[nev] = s;
                        }
                        ++nev;
                        cur = __dmul_rn(cur, s);
                        tmax = __dmul_rn(tmax, s);
                    }
                }
            }
#pragma unroll
            for (int b = 0; b < CW_B; ++b)
                if (i0 - b >= 1) raw[i0 - b - 1] = vb[b];
        }
        *S.maxv = tmax;
    } else {
        double* __restrict__ raw = S.raw;
        raw[0] = W.chi0;
        raw[1] = W.chi1;
        double cm = W.chi0, cc = W.chi1;
        double Qm = q_of(ell, r[0], U[0], k2), Qc = q_of(ell, r[1], U[1], k2);
        double rp = r[0], rc = r[1];
        for (int i0 = 1; i0 < m - 1; i0 += CW_B) {
          double rb[CW_B], ub[CW_B], vout[CW_B];
#pragma unroll
          for (int b = 0; b < CW_B; ++b) {
              const int i = i0 + b;
              rb[b] = i + 1 < m ? r[i + 1] : 0.0;
              ub[b] = i + 1 < m ? U[i + 1] : 0.0;
          }
#pragma unroll
          for (int b = 0; b < CW_B; ++b) {
            const int i = i0 + b;
            if (i < m - 1) {
            const double rn = rb[b];
            const double h1 = __dsub_rn(rc, rp), h2 = __dsub_rn(rn, rc);
            const double Qn = q_of(ell, rn, ub[b], k2);
            const double a_prev = __dsub_rn(1.0, __dmul_rn(__ddiv_rn(__dmul_rn(h1, h1), 12.0), Qm));
            const double b_curr = __dadd_rn(2.0, __dmul_rn(__ddiv_rn(__dmul_rn(5.0, __dmul_rn(h1, h2)), 6.0), Qc));
            double a_next = __dsub_rn(1.0, __dmul_rn(__ddiv_rn(__dmul_rn(h2, h2), 12.0), Qn));
            if (fabs(a_next) < 1e-15) a_next = a_next >= 0.0 ? 1e-15 : -1e-15;
            double v = __ddiv_rn(__dsub_rn(__dmul_rn(b_curr, cc), __dmul_rn(a_prev, cm)), a_next);
            vout[b] = v;
            cm = cc;
            cc = v;
            Qm = Qc;
            Qc = Qn;
            rp = rc;
            rc = rn;
            if ((i + 1) % W.renorm == 0) {
                const double mv = fmax(fabs(cm), fabs(cc));
                double s = 0.0;
                if (mv > 1e50) s = __ddiv_rn(1e50, mv);
                else if (mv < 1e-50 && mv > 0.0) s = __ddiv_rn(1.0, __dmul_rn(mv, 1e10));
                if (s != 0.0) {
                    if (nev < DWBA_CW_MAX_EVENTS) {
                        S.ev_pos[nev] = i + 1;      // applies to positions <= i+1
                        S.ev_scale[nev] = s;
                    }
                    ++nev;
                    cm = __dmul_rn(cm, s);
                    cc = __dmul_rn(cc, s);
                }
            }
                      }
          }
#pragma unroll
          for (int b = 0; b < CW_B; ++b)
              if (i0 + b < m - 1) raw[i0 + b + 1] = vout[b];
        }
        *S.maxv = 0.0;
    }
    *S.ev_count = nev;
    if (nev > DWBA_CW_MAX_EVENTS) over = true;
    status[w] = over ? 1 : 0;
}

// chi[w][j]: zero below start, the rescaled raw values from start on.
__global__ __launch_bounds__(CT) void cw_apply_kernel(int n_waves, const dwba_cw_wave* __restrict__ waves, int ld,
                                                      char* __restrict__ work, double* __restrict__ chi) {
    const int w = blockIdx.y;
    const dwba_cw_wave W = waves[w];
    CwWork S = work_of(work, w, ld);
    const int nev = min(*S.ev_count, DWBA_CW_MAX_EVENTS);
    double* __restrict__ out = chi + (size_t)w * ld;
    for (int j = blockIdx.x * CT + threadIdx.x; j < ld; j += gridDim.x * CT) {
        if (j >= W.n) { out[j] = 0.0; continue; }
        if (j < W.start) { out[j] = 0.0; continue; }
        const int q = j - W.start;
        double v = S.raw[q];
        if (W.method == DWBA_CW_JOHNSON) {
            for (int e = 0; e < nev; ++e)            // recorded in time order: decreasing position
                if (S.ev_pos[e] <= q) v = __dmul_rn(v, S.ev_scale[e]);
            const double mx = *S.maxv;
            if (mx > 0.0) v = __ddiv_rn(v, mx);
        } else {
            for (int e = 0; e < nev; ++e)            // increasing position
                if (S.ev_pos[e] >= q) v = __dmul_rn(v, S.ev_scale[e]);
        }
        out[j] = v;
    }
}

// Fornberg's first-derivative weights at x[c] from the 5 points x[0..4]
// (the recursion of continuum.py:957-979 for m = 1), dotted with f.
__device__ inline double fornberg5(const double* x, double x0, const double* f) {
    double c0[5], c1[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) { c0[i] = 0.0; c1[i] = 0.0; }
    c0[0] = 1.0;
    double cc1 = 1.0, c4 = __dsub_rn(x[0], x0);
#pragma unroll
    for (int i = 1; i < 5; ++i) {
        double c2 = 1.0;
        const double c5 = c4;
        c4 = __dsub_rn(x[i], x0);
#pragma unroll
        for (int j = 0; j < i; ++j) {
            const double c3 = __dsub_rn(x[i], x[j]);
            c2 = __dmul_rn(c2, c3);
            if (j == i - 1) {
                c1[i] = __ddiv_rn(__dmul_rn(cc1, __dsub_rn(__dmul_rn(1.0, c0[i - 1]), __dmul_rn(c5, c1[i - 1]))), c2);
                c0[i] = __ddiv_rn(-__dmul_rn(__dmul_rn(cc1, c5), c0[i - 1]), c2);
            }
            c1[j] = __ddiv_rn(__dsub_rn(__dmul_rn(c4, c1[j]), __dmul_rn(1.0, c0[j])), c3);
            c0[j] = __ddiv_rn(__dmul_rn(c4, c0[j]), c3);
        }
        cc1 = c2;
    }
    // np.dot of the 5 weights with f: pairwise order of a length-5 dot is sequential
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < 5; ++i) s = __dadd_rn(s, __dmul_rn(c1[i], f[i]));
    return s;
}

__device__ inline double deriv_at(const double* __restrict__ r, const double* __restrict__ f, int n, int idx) {
    const int s = min(max(idx - 2, 0), n - 5);
    double x[5], v[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) { x[i] = r[s + i]; v[i] = f[s + i]; }
    return fornberg5(x, r[idx], v);
}

// dchi[w][j] on the wave's propagated range [start, n) (the stencil stays in it), 0 below.
__global__ __launch_bounds__(CT) void cw_deriv_kernel(int n_waves, const dwba_cw_wave* __restrict__ waves, int ld,
                                                      const double* __restrict__ chi, double* __restrict__ dchi) {
    const int w = blockIdx.y;
    const dwba_cw_wave W = waves[w];
    const double* __restrict__ c = chi + (size_t)w * ld + W.start;
    const double* __restrict__ r = W.r + W.start;
    const int m = W.n - W.start;
    double* __restrict__ out = dchi + (size_t)w * ld;
    for (int j = blockIdx.x * CT + threadIdx.x; j < ld; j += gridDim.x * CT) {
        if (j < W.start || j >= W.n) { out[j] = 0.0; continue; }
        out[j] = m >= 5 ? deriv_at(r, c, m, j - W.start) : 0.0;
    }
}

// ---- spherical Bessel functions ------------------------------------------
__device__ inline double sph_j(int l, double x) {
    if (x == 0.0) return l == 0 ? 1.0 : 0.0;
    double s, c;
    sincos(x, &s, &c);
    const double j0 = s / x;
    if (l == 0) return j0;
    const double j1 = s / (x * x) - c / x;
    if (l == 1) return j1;
    if (x > (double)l) {
        double jm = j0, jc = j1;
        for (int n = 1; n < l; ++n) {
            const double jn = (2.0 * n + 1.0) / x * jc - jm;
            jm = jc;
            jc = jn;
        }
        return jc;
    }
    // Miller: f_{n-1} = (2n+1)/x f_n - f_{n+1} from far above l, normalised at the end
    const int top = l + 32 + (int)(2.0 * sqrt((double)l + 1.0)) + (int)x;
    double fp = 0.0, f = 1e-290, fl = 0.0, f1 = 0.0;
    for (int n = top; n >= 1; --n) {
        const double fm = (2.0 * n + 1.0) / x * f - fp;
        fp = f;
        f = fm;
        if (n - 1 == l) fl = f;
        if (n - 1 == 1) f1 = f;
        if (fabs(f) > 1e250) {
            f *= 1e-250;
            fp *= 1e-250;
            fl *= 1e-250;
            f1 *= 1e-250;
        }
    }
    // f = f_0; scale by j_0 or j_1, whichever is further from a zero
    if (fabs(j0) >= fabs(j1)) return fl * (j0 / f);
    return fl * (j1 / f1);
}

__device__ inline double sph_y(int l, double x) {
    double s, c;
    sincos(x, &s, &c);
    double ym = -c / x;
    if (l == 0) return ym;
    double yc = -c / (x * x) - s / x;
    for (int n = 1; n < l; ++n) {
        const double yn = (2.0 * n + 1.0) / x * yc - ym;
        ym = yc;
        yc = yn;
    }
    return yc;
}

// Asymptotic wave at r (continuum.py:1072-1162): neutral A (rho j_l cos d - rho y_l sin d),
// ionic A (sin th cos d - cos th sin d), th = rho + eta ln 2 rho - l pi/2 + sigma.
__device__ inline double asym_wave(const dwba_cw_finish& F, double cd, double sd, double r) {
    const double rho = F.k * r;
    if (rho < 1e-10) return 0.0;
    const double A = sqrt(2.0 / PI_C);
    if (!F.coulomb) {
        double yl = sph_y(F.l, rho);
        if (!isfinite(yl) || !(fabs(yl) < 1e50)) yl = -1e30;
        return A * (rho * sph_j(F.l, rho) * cd - rho * yl * sd);
    }
    const double th = rho + F.eta * log(2.0 * rho) - (F.l * PI_C / 2.0) + F.sigma;
    double s, c;
    sincos(th, &s, &c);
    return A * (s * cd - c * sd);
}

template <int NR>
__device__ inline void block_sums(double (&v)[NR], double (*s_red)[CT / 64]) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NR; ++q) {
        const double t = wave_sum(v[q]);
        if (lane == 0) s_red[q][wv] = t;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NR; ++q) {
        double t = 0.0;
        for (int x = 0; x < CT / 64; ++x) t += s_red[q][x];
        v[q] = t;
    }
    __syncthreads();
}

// Tail least squares: sums over j in [tail, n) of s^2, s c, c^2, chi s, chi c and max |chi|.
__global__ __launch_bounds__(CT) void cw_tail_sums_kernel(int n_waves, const dwba_cw_finish* __restrict__ fin, int ld,
                                                          const double* __restrict__ chi, double* __restrict__ sums) {
    __shared__ double s_red[6][CT / 64];
    const int w = blockIdx.x;
    const dwba_cw_finish F = fin[w];
    const double* __restrict__ c = chi + (size_t)w * ld;
    double v[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    double mx = 0.0;
    for (int j = F.tail + threadIdx.x; j < F.n; j += CT) {
        const double rr = F.r[j];
        double ph;
        if (!F.coulomb) {
            ph = F.k * rr - (double)F.l * PI_C / 2.0;
        } else {
            const double rho = F.k * rr;
            ph = rho + F.eta * log(2.0 * rho) - (F.l * PI_C / 2.0) + F.sigma;
        }
        double s, co;
        sincos(ph, &s, &co);
        const double y = c[j];
        v[0] = fma(s, s, v[0]);
        v[1] = fma(s, co, v[1]);
        v[2] = fma(co, co, v[2]);
        v[3] = fma(y, s, v[3]);
        v[4] = fma(y, co, v[4]);
        mx = fmax(mx, fabs(y));
    }
    block_sums<6>(v, s_red);
    // max over the block
    {
        double t = mx;
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) t = fmax(t, __shfl_xor(t, m, 64));
        if ((threadIdx.x & 63) == 0) s_red[0][threadIdx.x >> 6] = t;
        __syncthreads();
        if (threadIdx.x == 0) {
            double b = 0.0;
            for (int x = 0; x < CT / 64; ++x) b = fmax(b, s_red[0][x]);
            double* o = sums + (size_t)w * 6;
            for (int q = 0; q < 5; ++q) o[q] = v[q];
            o[5] = b;
        }
    }
}

// Stitching and amplitude (one workgroup per wave): chi is read at [0, idx_match]
// and the final wave written to out; scratch holds the asymptotic wave.
__global__ __launch_bounds__(CT) void cw_stitch_kernel(int n_waves, const dwba_cw_finish* __restrict__ fin, int ld,
                                                       const double* __restrict__ chi, const double* __restrict__ dchi,
                                                       double* __restrict__ scratch, double* __restrict__ out) {
    __shared__ double s_red[2][CT / 64];
    __shared__ double s_scale, s_div;
    const int w = blockIdx.x;
    const dwba_cw_finish F = fin[w];
    const double* __restrict__ c = chi + (size_t)w * ld;
    double* __restrict__ as = scratch + (size_t)w * ld;
    double* __restrict__ o = out + (size_t)w * ld;
    const int im = F.idx_match;
    const double cd = cos(F.delta), sd = sin(F.delta);
    const int j0 = max(im - 2, 0);
    for (int j = j0 + threadIdx.x; j < F.n; j += CT) as[j] = asym_wave(F, cd, sd, F.r[j]);
    __syncthreads();
    if (threadIdx.x == 0) {
        const double cm = c[im], dm = dchi[(size_t)w * ld + im];
        double scale = 1.0;
        if (fabs(cm) > 1e-100) {
            const double das = deriv_at(F.r, as, F.n, im);
            const double num = as[im] * cm + das * dm;
            const double den = cm * cm + dm * dm;
            scale = fabs(den) > 1e-100 ? num / den : as[im] / cm;
        }
        s_scale = scale;
    }
    __syncthreads();
    const double scale = s_scale;
    double v[2] = {0.0, 0.0};
    for (int j = threadIdx.x; j < ld; j += CT) {
        double x = 0.0;
        if (j < F.n) x = j <= im ? scale * c[j] : as[j];
        o[j] = x;
        if (j >= im && j < F.n) {
            const double wt = F.w_trapz[j];
            v[0] = fma(wt, x * x, v[0]);
            v[1] += wt;
        }
    }
    block_sums<2>(v, s_red);
    if (threadIdx.x == 0) {
        double div = 1.0;
        if (F.n - im > 10 && v[1] > 0.0) {
            const double amp = sqrt(v[0] / v[1]) * sqrt(2.0);
            if (fabs(amp - 1.0) > 0.05 && amp > 1e-50) div = amp;
        }
        s_div = div;
    }
    __syncthreads();
    const double div = s_div;
    if (div != 1.0)
        for (int j = threadIdx.x; j < F.n; j += CT) o[j] = o[j] / div;
}

// Analytic bypass: chi = sqrt(2/pi) (r j_l(k r) k), delta = clip(-k sum w U j_l^2 r^2).
__global__ __launch_bounds__(CT) void cw_bypass_kernel(int n_waves, const dwba_cw_bypass* __restrict__ byp, int ld,
                                                       double* __restrict__ out, double* __restrict__ delta) {
    __shared__ double s_red[1][CT / 64];
    const int w = blockIdx.x;
    const dwba_cw_bypass B = byp[w];
    double* __restrict__ o = out + (size_t)w * ld;
    const double A = sqrt(2.0 / PI_C);
    double v[1] = {0.0};
    for (int j = threadIdx.x; j < ld; j += CT) {
        if (j >= B.n) { o[j] = 0.0; continue; }
        const double r = B.r[j];
        const double jl = sph_j(B.l, B.k * r);
        o[j] = A * ((r * jl) * B.k);
        v[0] = fma(B.w_trapz[j], (B.U[j] * (jl * jl)) * (r * r), v[0]);
    }
    block_sums<1>(v, s_red);
    if (threadIdx.x == 0) delta[w] = fmin(fmax(-B.k * v[0], -PI_C / 2.0), PI_C / 2.0);
}

}  // namespace
}  // namespace dwba

using namespace dwba;

extern "C" {

size_t dwba_cw_workspace_bytes(int32_t n_waves, int32_t ld) {
    if (n_waves < 1 || ld < 8) return 0;
    return (size_t)n_waves * cw_per_wave(ld);
}

int32_t dwba_cw_propagate_f64(int32_t n_waves, const dwba_cw_wave* d_waves, int32_t ld, double* d_chi, double* d_dchi,
                              void* d_work, size_t work_bytes, int32_t* d_status, void* stream) {
    DWBA_REQUIRE(initialised(), "dwba_cw_propagate_f64: call dwba_init first");
    DWBA_REQUIRE(n_waves >= 1 && n_waves <= 65535 && ld >= 8 && d_waves && d_chi && d_dchi && d_work && d_status,
                 "dwba_cw_propagate_f64: bad arguments");
    if (work_bytes < dwba_cw_workspace_bytes(n_waves, ld))
        return set_error(DWBA_ERR_WORKSPACE, "dwba_cw_propagate_f64: workspace %zu < %zu bytes", work_bytes,
                         dwba_cw_workspace_bytes(n_waves, ld));
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    char* wk = reinterpret_cast<char*>(d_work);
    hipLaunchKernelGGL(cw_propagate_kernel, dim3((n_waves + 63) / 64), dim3(64), 0, s, n_waves, d_waves, ld, wk,
                       d_status);
    DWBA_LAUNCH_CHECK("cw_propagate_kernel");
    const unsigned gx = (unsigned)std::min((ld + CT - 1) / CT, 64);
    hipLaunchKernelGGL(cw_apply_kernel, dim3(gx, n_waves), dim3(CT), 0, s, n_waves, d_waves, ld, wk, d_chi);
    DWBA_LAUNCH_CHECK("cw_apply_kernel");
    hipLaunchKernelGGL(cw_deriv_kernel, dim3(gx, n_waves), dim3(CT), 0, s, n_waves, d_waves, ld, d_chi, d_dchi);
    DWBA_LAUNCH_CHECK("cw_deriv_kernel");
    return DWBA_OK;
}

int32_t dwba_cw_tail_sums_f64(int32_t n_waves, const dwba_cw_finish* d_fin, int32_t ld, const double* d_chi,
                              double* d_sums, void* stream) {
    DWBA_REQUIRE(initialised(), "dwba_cw_tail_sums_f64: call dwba_init first");
    DWBA_REQUIRE(n_waves >= 1 && ld >= 8 && d_fin && d_chi && d_sums, "dwba_cw_tail_sums_f64: bad arguments");
    hipLaunchKernelGGL(cw_tail_sums_kernel, dim3(n_waves), dim3(CT), 0, reinterpret_cast<hipStream_t>(stream), n_waves,
                       d_fin, ld, d_chi, d_sums);
    DWBA_LAUNCH_CHECK("cw_tail_sums_kernel");
    return DWBA_OK;
}

int32_t dwba_cw_stitch_f64(int32_t n_waves, const dwba_cw_finish* d_fin, int32_t ld, const double* d_chi,
                           const double* d_dchi, double* d_scratch, double* d_out, void* stream) {
    DWBA_REQUIRE(initialised(), "dwba_cw_stitch_f64: call dwba_init first");
    DWBA_REQUIRE(n_waves >= 1 && ld >= 8 && d_fin && d_chi && d_dchi && d_scratch && d_out,
                 "dwba_cw_stitch_f64: bad arguments");
    hipLaunchKernelGGL(cw_stitch_kernel, dim3(n_waves), dim3(CT), 0, reinterpret_cast<hipStream_t>(stream), n_waves,
                       d_fin, ld, d_chi, d_dchi, d_scratch, d_out);
    DWBA_LAUNCH_CHECK("cw_stitch_kernel");
    return DWBA_OK;
}

int32_t dwba_cw_bypass_f64(int32_t n_waves, const dwba_cw_bypass* d_byp, int32_t ld, double* d_out, double* d_delta,
                           void* stream) {
    DWBA_REQUIRE(initialised(), "dwba_cw_bypass_f64: call dwba_init first");
    DWBA_REQUIRE(n_waves >= 1 && ld >= 8 && d_byp && d_out && d_delta, "dwba_cw_bypass_f64: bad arguments");
    hipLaunchKernelGGL(cw_bypass_kernel, dim3(n_waves), dim3(CT), 0, reinterpret_cast<hipStream_t>(stream), n_waves,
                       d_byp, ld, d_out, d_delta);
    DWBA_LAUNCH_CHECK("cw_bypass_kernel");
    return DWBA_OK;
}

}  // extern "C"
