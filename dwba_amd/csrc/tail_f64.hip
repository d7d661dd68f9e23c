// FP64 outer tail of the direct integral, one wavefront per (pair, L).
//
// Restates, on the GPU, the CPU chain radial_ME_all_L uses for L >= 1
// excitation tails (dwba_matrix_elements.py:1165-1192, :1216-1241, and the
// legacy analytic tail :1257-1268):
//
//   I_out = 1/2 [Re J(Phi_-) - Re J(Phi_+)] + T(r_max)        (advanced/full_split)
//   I_out = T(r_match)                                        (legacy)
//   J(Phi) = int_{r_m}^{r_max} M_L r^-(L+1) exp(i Phi(r)) dr
//
// J is the selector compute_outer_integral_oscillatory
// (oscillatory_integrals.py:1361-1451): Filon with the constant frequency
// |Phi'(mid)| (filon_oscillatory_integral :1280-1354, _filon_segment_complex
// :1165-1277, segment cap DWBA_MAX_FILON_SEGMENTS and node up-scaling past
// DWBA_MAX_EFFECTIVE_DPHI) when |Phi''(mid)| h^2 < 0.1, else Levin collocation
// on 8 Chebyshev-Lobatto nodes per segment (:1024-1158).  Phases are the
// sinA*sinB phases of _get_dwba_outer_phase_functions (:350-456); T is
// _analytical_multipole_tail (:2064-2184).  The reference sums segment
// contributions with math.fsum; here each lane keeps a double-double
// (TwoSum) accumulator (Levin path) or an FP64 partial sum (Filon path), and
// the lanes are combined in double-double; the result agrees with math.fsum
// to ~1e-15 of the segment size.
//
// Lanes of the wavefront take segments: contiguous runs on the all-L Filon
// path, round-robin on the per-L Levin path.
#include "common.hpp"

#include <type_traits>

namespace dwba {
namespace {

constexpr double PI = 3.14159265358979323846;

struct DD {
    double hi, lo;
};

__device__ inline DD dd_add(DD a, double b) {
    const double s = a.hi + b;
    const double bb = s - a.hi;
    const double e = (a.hi - (s - bb)) + (b - bb);
    const double lo = a.lo + e;
    const double hi = s + lo;
    return {hi, lo - (hi - s)};
}

__device__ inline DD dd_add(DD a, DD b) {
    const double s = a.hi + b.hi;
    const double bb = s - a.hi;
    const double e = (a.hi - (s - bb)) + (b.hi - bb);
    const double lo = (a.lo + b.lo) + e;
    const double hi = s + lo;
    return {hi, lo - (hi - s)};
}

__device__ inline DD dd_wave_sum(DD v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        DD o{__shfl_xor(v.hi, m, 64), __shfl_xor(v.lo, m, 64)};
        v = dd_add(v, o);
    }
    return v;
}

// Per-L sums of the 64 lanes' partial sums acc[L], L < LE, combined in
// double-double: 8 multipoles at a time pass through a 64 x 9 LDS tile (a
// small tile keeps the LDS per wavefront low, so it does not bound the
// number of resident waves); lane (g, c) adds rows 8g..8g+7 of column c, and
// the 8 row groups are joined by xor shuffles (TwoSum's error term is exact,
// so both lanes of a shuffle step hold the same sum).  out[L] = M_L * sum for
// 1 <= L <= Lm with |M_L| > 1e-12, 0 otherwise (L < NL).  One wavefront per
// workgroup.
constexpr int SUM_COLS = 8;
template <int NL, int LE>
__device__ __forceinline__ void lane_sums_out(const double* acc, double (*s_acc)[SUM_COLS + 1], int lane,
                                              const double* moments, int Lm, double* out) {
    const int c = lane & (SUM_COLS - 1), g = lane / SUM_COLS;
#pragma unroll
    for (int L0 = 0; L0 < NL; L0 += SUM_COLS) {
        if (L0 > 0) __syncthreads();              // the previous tile has been read
#pragma unroll
        for (int q = 0; q < SUM_COLS; ++q) s_acc[lane][q] = (L0 + q < LE) ? acc[L0 + q] : 0.0;
        __syncthreads();
        DD t{0.0, 0.0};
#pragma unroll
        for (int k = 0; k < 64 / SUM_COLS; ++k) t = dd_add(t, s_acc[g * (64 / SUM_COLS) + k][c]);
#pragma unroll
        for (int m = SUM_COLS; m < 64; m <<= 1) {
            DD o{__shfl_xor(t.hi, m, 64), __shfl_xor(t.lo, m, 64)};
            t = dd_add(t, o);
        }
        const int L = L0 + c;
        if (g == 0 && L < NL) {
            const double M = (L >= 1 && L <= Lm && L < LE) ? moments[L] : 0.0;
            out[L] = fabs(M) > 1e-12 ? M * (t.hi + t.lo) : 0.0;
        }
    }
}

struct cplx {
    double re, im;
};
__device__ inline cplx cadd(cplx a, cplx b) { return {a.re + b.re, a.im + b.im}; }
__device__ inline cplx csub(cplx a, cplx b) { return {a.re - b.re, a.im - b.im}; }
__device__ inline cplx cmul(cplx a, cplx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ inline cplx cscale(cplx a, double s) { return {a.re * s, a.im * s}; }
__device__ inline cplx cexpi(double phi) {
    double s, c;
    sincos(phi, &s, &c);
    return {c, s};
}
__device__ inline cplx cdiv(cplx a, cplx b) {
    // Smith's algorithm
    if (fabs(b.re) >= fabs(b.im)) {
        const double r = b.im / b.re, d = b.re + b.im * r;
        return {(a.re + a.im * r) / d, (a.im - a.re * r) / d};
    }
    const double r = b.re / b.im, d = b.re * r + b.im;
    return {(a.re * r + a.im) / d, (a.im * r - a.re) / d};
}

struct Phase {
    double ka, kb, ea, eb, la, lb;  // l as double
    bool use_ea, use_eb, use_la, use_lb;
    double kmin, kplus, cmin, cplus;
};

__device__ inline double phi(const Phase& P, int sgn, double r) {
    double res;
    if (sgn < 0) {
        res = P.kmin * r + P.cmin;
        if (P.use_ea) res += P.ea * log(2 * P.ka * r + 1e-30);
        if (P.use_eb) res -= P.eb * log(2 * P.kb * r + 1e-30);
        if (P.use_la) res -= (P.la * (P.la + 1)) / (2.0 * P.ka * r);
        if (P.use_lb) res += (P.lb * (P.lb + 1)) / (2.0 * P.kb * r);
    } else {
        res = P.kplus * r + P.cplus;
        if (P.use_ea) res += P.ea * log(2 * P.ka * r + 1e-30);
        if (P.use_eb) res += P.eb * log(2 * P.kb * r + 1e-30);
        if (P.use_la) res -= (P.la * (P.la + 1)) / (2.0 * P.ka * r);
        if (P.use_lb) res -= (P.lb * (P.lb + 1)) / (2.0 * P.kb * r);
    }
    return res;
}

__device__ inline double dphi(const Phase& P, int sgn, double r) {
    double res = sgn < 0 ? P.kmin : P.kplus;
    if (P.use_ea) res += P.ea / r;
    if (P.use_eb) res = sgn < 0 ? res - P.eb / r : res + P.eb / r;
    if (P.use_la) res += (P.la * (P.la + 1)) / (2.0 * P.ka * (r * r));
    if (P.use_lb) {
        const double t = (P.lb * (P.lb + 1)) / (2.0 * P.kb * (r * r));
        res = sgn < 0 ? res - t : res + t;
    }
    return res;
}

__device__ inline double d2phi(const Phase& P, int sgn, double r) {
    double res = 0.0;
    if (P.use_ea) res -= P.ea / (r * r);
    if (P.use_eb) res = sgn < 0 ? res + P.eb / (r * r) : res - P.eb / (r * r);
    if (P.use_la) res -= (P.la * (P.la + 1)) / (P.ka * pow(r, 3.0));
    if (P.use_lb) {
        const double t = (P.lb * (P.lb + 1)) / (P.kb * pow(r, 3.0));
        res = sgn < 0 ? res + t : res - t;
    }
    return res;
}

struct Env {
    double M;   // moment
    int n1;     // L + 1
};
__device__ inline double envelope(const Env& E, double r) { return r > 1e-6 ? E.M / pow(r, (double)E.n1) : 0.0; }

// 1/x for 1e-6 < x < 1e300: the hardware reciprocal estimate refined by two
// Newton steps (each at least doubles its correct bits), within an ulp or so of
// the correctly rounded quotient at a third of the division's instructions.
__device__ __forceinline__ double recip(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

struct LevinTab {
    double D[8][8];   // Chebyshev differentiation matrix on [-1, 1] (descending Lobatto points)
    double xr[8];     // ascending Lobatto points
};

// _levin_segment_complex for every multipole at once.  The collocation
// matrix A = D 2/h + i diag(Phi'(x_k)) does not depend on L; only the
// right-hand side f_k = M_L x_k^-(L+1) does, and the segment value
// u_7 e^{i Phi(x_7)} - u_0 e^{i Phi(x_0)} with u = A^-1 f is g . f,
// g = e^{i Phi(x_7)} (row 7 of A^-1) - e^{i Phi(x_0)} (row 0 of A^-1).  The two
// rows solve A^T y = e_7 and A^T y = e_0 (one LU with partial pivoting on
// |re| + |im|, LAPACK's izamax convention).  The reference solves A u = f per L;
// both are backward stable, so they differ by about cond(A) 2^-53 |J|, the
// noise the reference's own solver carries.  Returns Re g and 1/x_k (0 where
// the envelope is cut off, x <= 1e-6); false for a degenerate segment (value 0).
__device__ bool levin_weights(const Phase& P, int sgn, double a, double b, const LevinTab& T, double* gre,
                              double* ix) {
    double x[8];
    for (int k = 0; k < 8; ++k) x[k] = 0.5 * (b - a) * (T.xr[k] + 1) + a;
    const double h = x[7] - x[0];
    if (h < 1e-15) return false;
    const double sc = 2.0 / h;
    cplx M[8][8];                 // A^T
    cplx y0[8], y7[8];
    for (int i = 0; i < 8; ++i) {
        for (int j = 0; j < 8; ++j) M[i][j] = {T.D[j][i] * sc, 0.0};
        M[i][i].im = dphi(P, sgn, x[i]);
        y0[i] = {i == 0 ? 1.0 : 0.0, 0.0};
        y7[i] = {i == 7 ? 1.0 : 0.0, 0.0};
    }
    for (int c = 0; c < 8; ++c) {
        int p = c;
        double best = fabs(M[c][c].re) + fabs(M[c][c].im);
        for (int i = c + 1; i < 8; ++i) {
            const double v = fabs(M[i][c].re) + fabs(M[i][c].im);
            if (v > best) { best = v; p = i; }
        }
        if (p != c) {
            for (int j = 0; j < 8; ++j) { const cplx t = M[c][j]; M[c][j] = M[p][j]; M[p][j] = t; }
            cplx t = y0[c]; y0[c] = y0[p]; y0[p] = t;
            t = y7[c]; y7[c] = y7[p]; y7[p] = t;
        }
        const cplx piv = M[c][c];
        if (piv.re == 0.0 && piv.im == 0.0) continue;
        for (int i = c + 1; i < 8; ++i) {
            const cplx l = cdiv(M[i][c], piv);
            for (int j = c + 1; j < 8; ++j) M[i][j] = csub(M[i][j], cmul(l, M[c][j]));
            y0[i] = csub(y0[i], cmul(l, y0[c]));
            y7[i] = csub(y7[i], cmul(l, y7[c]));
        }
    }
    for (int i = 7; i >= 0; --i) {
        cplx s0 = y0[i], s7 = y7[i];
        for (int j = i + 1; j < 8; ++j) {
            s0 = csub(s0, cmul(M[i][j], y0[j]));
            s7 = csub(s7, cmul(M[i][j], y7[j]));
        }
        y0[i] = cdiv(s0, M[i][i]);
        y7[i] = cdiv(s7, M[i][i]);
    }
    const cplx e0 = cexpi(phi(P, sgn, x[0])), e7 = cexpi(phi(P, sgn, x[7]));
    for (int k = 0; k < 8; ++k) {
        gre[k] = (e7.re * y7[k].re - e7.im * y7[k].im) - (e0.re * y0[k].re - e0.im * y0[k].im);
        ix[k] = x[k] > 1e-6 ? recip(x[k]) : 0.0;
    }
    return true;
}

// non-oscillatory branch (scipy quad in the reference): composite 8-point
// Gauss-Legendre on 512 panels, lanes over panels
__device__ cplx gauss_quad(const Env& E, const Phase& P, int sgn, double a, double b, bool const_phase, double p0,
                           int lane) {
    const double gx[4] = {0.1834346424956498, 0.5255324099163290, 0.7966664774136267, 0.9602898564975363};
    const double gw[4] = {0.3626837833783620, 0.3137066458778873, 0.2223810344533745, 0.1012285362903763};
    const int NP = 512;
    DD re{0, 0}, im{0, 0};
    for (int pnl = lane; pnl < NP; pnl += 64) {
        const double pa = a + (b - a) * pnl / NP, pb = a + (b - a) * (pnl + 1) / NP;
        const double c = 0.5 * (pa + pb), hw = 0.5 * (pb - pa);
        for (int q = 0; q < 8; ++q) {
            const double xx = c + (q < 4 ? -gx[3 - q] : gx[q - 4]) * hw;
            const double ww = (q < 4 ? gw[3 - q] : gw[q - 4]) * hw;
            const double ph = const_phase ? p0 : phi(P, sgn, xx);
            const double fv = envelope(E, xx) * ww;
            re = dd_add(re, fv * cos(ph));
            im = dd_add(im, fv * sin(ph));
        }
    }
    re = dd_wave_sum(re);
    im = dd_wave_sum(im);
    return {re.hi, im.hi};
}

struct TailParams {
    int max_filon_segments;
    double max_effective_dphi;
    LevinTab lev;
};

__device__ double analytic_tail(double r_m, const dwba_tail_job& J, int L, double overlap) {
    if (r_m < 1e-3 || fabs(overlap) < 1e-12) return 0.0;
    const double lt_i = fabs(J.eta_a) > 1e-10 ? J.eta_a * log(2 * J.k_a * r_m + 1e-30) : 0.0;
    const double lt_f = fabs(J.eta_b) > 1e-10 ? J.eta_b * log(2 * J.k_b * r_m + 1e-30) : 0.0;
    const double kd = J.k_a - J.k_b, ks = J.k_a + J.k_b;
    const int n = L + 1;
    double tail = 0.0;
    const double pic = -J.l_a * PI / 2 + J.sigma_a + J.delta_a + lt_i;
    const double pfc = -J.l_b * PI / 2 + J.sigma_b + J.delta_b + lt_f;
    const double pd = kd * r_m + pic - pfc;
    if (fabs(kd) > 1e-6) {
        const double t1 = sin(pd) / (kd * pow(r_m, (double)n));
        const double t1c = n * cos(pd) / ((kd * kd) * pow(r_m, (double)(n + 1)));
        tail += 0.5 * (t1 + t1c);
    } else if (n > 1) {
        tail += 0.5 * cos(pd) / ((n - 1) * pow(r_m, (double)(n - 1)));
    }
    const double ps = ks * r_m + pic + pfc;
    if (ks > 1e-6) {
        const double t2 = sin(ps) / (ks * pow(r_m, (double)n));
        const double t2c = n * cos(ps) / ((ks * ks) * pow(r_m, (double)(n + 1)));
        tail -= 0.5 * (t2 + t2c);
    }
    if (!isfinite(tail)) return 0.0;
    return tail * overlap;
}

// x^-(L+1) sums of one node for L = 1 .. LE-1 (ix = 1/x, 0 where the
// envelope is cut off, x <= 1e-6); M_L multiplies the whole sum at the end.
template <int LE>
__device__ __forceinline__ void sweep1(double* acc, double w0, double ix0) {
    double p0x = ix0 * ix0;
#pragma unroll
    for (int L = 1; L < LE; ++L) {
        acc[L] = fma(w0, p0x, acc[L]);
        p0x *= ix0;
    }
}

// The same for two nodes.
template <int LE>
__device__ __forceinline__ void sweep2(double* acc, double w0, double ix0, double w1, double ix1) {
    double p0x = ix0 * ix0, p1x = ix1 * ix1;
#pragma unroll
    for (int L = 1; L < LE; ++L) {
        acc[L] = fma(w0, p0x, acc[L]);
        acc[L] = fma(w1, p1x, acc[L]);
        p0x *= ix0;
        p1x *= ix1;
    }
}

// The same for four nodes, added to each acc[L] in node order.  The powers
// x^-(L+1) = x^-(2+c) (x^-4)^a (L = 1 + 4a + c) come in blocks of four: per
// node four products w x^-(2+c) and one power of x^-4 per block feed four
// FMAs, about 1.4 instructions per node and multipole instead of an FMA and
// a multiply each (the products round differently from a running power, at
// the 1e-16 level of each term).
template <int LE>
__device__ __forceinline__ void sweep4(double* acc, double w0, double ix0, double w1, double ix1, double w2,
                                       double ix2, double w3, double ix3) {
    const double ws[4] = {w0, w1, w2, w3}, ys[4] = {ix0, ix1, ix2, ix3};
    double u[4][4], z[4], za[4];
#pragma unroll
    for (int n = 0; n < 4; ++n) {
        const double y2 = ys[n] * ys[n];
        u[n][0] = ws[n] * y2;
        u[n][1] = u[n][0] * ys[n];
        u[n][2] = u[n][1] * ys[n];
        u[n][3] = u[n][2] * ys[n];
        z[n] = y2 * y2;
        za[n] = z[n];
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int L = 1 + c;
        if (L < LE) acc[L] = (((acc[L] + u[0][c]) + u[1][c]) + u[2][c]) + u[3][c];
    }
#pragma unroll
    for (int a = 1; 1 + 4 * a < LE; ++a) {
        if (a > 1) {
#pragma unroll
            for (int n = 0; n < 4; ++n) za[n] *= z[n];
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int L = 1 + 4 * a + c;
            if (L < LE) {
                acc[L] = fma(u[0][c], za[0], acc[L]);
                acc[L] = fma(u[1][c], za[1], acc[L]);
                acc[L] = fma(u[2][c], za[2], acc[L]);
                acc[L] = fma(u[3][c], za[3], acc[L]);
            }
        }
    }
}

// The same for two nodes in the block form of sweep4: sweep2b on nodes 0, 1 and
// then on nodes 2, 3 makes exactly sweep4's operations in its order, with half
// its products live at a time.
template <int LE>
__device__ __forceinline__ void sweep2b(double* acc, double w0, double ix0, double w1, double ix1) {
    const double ws[2] = {w0, w1}, ys[2] = {ix0, ix1};
    double u[2][4], z[2], za[2];
#pragma unroll
    for (int n = 0; n < 2; ++n) {
        const double y2 = ys[n] * ys[n];
        u[n][0] = ws[n] * y2;
        u[n][1] = u[n][0] * ys[n];
        u[n][2] = u[n][1] * ys[n];
        u[n][3] = u[n][2] * ys[n];
        z[n] = y2 * y2;
        za[n] = z[n];
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int L = 1 + c;
        if (L < LE) acc[L] = (acc[L] + u[0][c]) + u[1][c];
    }
#pragma unroll
    for (int a = 1; 1 + 4 * a < LE; ++a) {
        if (a > 1) {
#pragma unroll
            for (int n = 0; n < 2; ++n) za[n] *= z[n];
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int L = 1 + 4 * a + c;
            if (L < LE) {
                acc[L] = fma(u[0][c], za[0], acc[L]);
                acc[L] = fma(u[1][c], za[1], acc[L]);
            }
        }
    }
}

// sin, cos of theta and the Filon coefficients alpha, beta, gamma at theta
// (_filon_segment_complex's formulas, Taylor series below |theta| = 0.3).
__device__ __forceinline__ void filon_coeffs(double th, double& rs, double& rc, double& al, double& be, double& ga) {
    sincos(th, &rs, &rc);
    if (fabs(th) < 0.3) {
        const double t2 = th * th, t3 = th * th * th;
        al = 2 * t3 / 45 * (1 - t2 / 7 + t2 * t2 / 189);
        be = 2.0 / 3 + 2 * t2 / 15 - 4 * (t2 * t2) / 105 + 2 * (t2 * t2 * t2) / 567;
        ga = 4.0 / 3 - 2 * t2 / 15 + 1 * (t2 * t2) / 210 - (t2 * t2 * t2) / 11340;
    } else {
        const double sd = 2.0 * rs * rc, cd = 1.0 - 2.0 * rs * rs;   // sin, cos of 2 theta
        const double t2 = th * th, it3 = 1.0 / (th * th * th);
        al = (t2 + th * sd / 2 - 2 * (rs * rs)) * it3;
        be = 2 * (th * (1 + cd) - sd) * it3;
        ga = 4 * (rs - th * rc) * it3;
    }
}

// Node weights (wa, wm, wb) of one triple x0 < x1 < x2 with the phases
// (s_k, c_k) = (sin, cos)(omega x_k + p0): the real part of the Filon sum, or
// simpson_odd times the cosine of the mid phase where the triple (or, with
// `small`, its whole segment, mid phase cosine cmid) does not oscillate.
__device__ __forceinline__ void triple_weights(double x0, double x1, double x2, double s0, double c0, double c1,
                                               double s2, double c2, bool small, double cmid, double om, double p0,
                                               double inv_om, double al, double be, double ga, double& wa,
                                               double& wm, double& wb) {
    const double h3 = x2 - x0;
    if (small || fabs(om * h3) < 1e-3 || h3 < 1e-15) {
        const double wscale = small ? cmid : (h3 < 1e-15 ? 0.0 : cos(om * (x0 + x2) / 2 + p0));
        const double h0 = x1 - x0, h1 = x2 - x1;
        const double hsum = h0 + h1, hprod = h0 * h1;
        const double h0divh1 = h1 != 0.0 ? h0 / h1 : 0.0;
        const double q = hsum / 6.0;
        wa = q * (2.0 - (h0divh1 != 0.0 ? 1.0 / h0divh1 : 0.0)) * wscale;
        wm = q * (hsum * (hprod != 0.0 ? hsum / hprod : 0.0)) * wscale;
        wb = q * (2.0 - h0divh1) * wscale;
    } else {
        // Re[(h/2)(al (-i/om)(f2 e2 - f0 e0) + be (f0 e0 + f2 e2) + ga f1 e1)]
        const double hh = h3 / 2, cA = al * inv_om;
        wa = hh * (cA * s0 + be * c0);
        wb = hh * (-cA * s2 + be * c2);
        wm = hh * (ga * c1);
    }
}

// Filon segments [s_begin, s_end) of one (pair, sign), multipoles L < LE.
// Each lane takes a contiguous run of segments; a segment of neff (odd)
// equally spaced nodes is walked triple by triple (the recursion of
// _filon_segment_complex).  Per triple the real part of the Filon sum is
// wa f(x0) + wm f(x1) + wb f(x2), so a node shared by two triples (or two
// segments) is swept once with the sum of its weights.  The Filon
// coefficients depend on theta = omega * step only and are formed once (the
// last segment apart); the node phases exp(i (omega x + p0)) advance by the
// rotation exp(i theta) from the segment start (the reference evaluates each
// node phase directly: the two agree to a few ulp of the phase).  Each lane
// adds to per-L FP64 partial sums (a few hundred terms of similar size); the
// lanes are combined in double-double.  The reference sums segments with
// math.fsum; the two agree to ~1e-15 of the segment size.
// A wavefront-uniform value moved to scalar registers (every lane holds the
// same value; the vector registers it frees let three wavefronts share a SIMD).
__device__ __forceinline__ double uni(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readfirstlane((int)b), hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
template <int LE>
__device__ __forceinline__ void filon_sweep(int neff, double om, double p0, double r_m, double r_max, double bstep,
                                            long nseg, long s_begin, long s_end, int lane, double* acc) {
    om = uni(om); p0 = uni(p0); r_m = uni(r_m); r_max = uni(r_max); bstep = uni(bstep);
    const int nt = (neff - 1) / 2;
    const double inv_om = uni(-1.0 / om);
    const long per = (s_end - s_begin + 63) / 64;
    const long s_lo = s_begin + (long)lane * per;
    const long s_hi = s_lo + per < s_end ? s_lo + per : s_end;
    double w_pend = 0.0, ix_pend = 0.0;           // weight still to sweep at the lane's current node
    // Segments are [r_m + s bstep, r_m + (s+1) bstep], so theta = omega * step
    // is one value up to rounding: its sine, cosine and Filon coefficients are
    // formed once; only the last segment (ending at r_max) gets its own.
    double rs_u, rc_u, al_u, be_u, ga_u;
    filon_coeffs(om * (bstep / (neff - 1)), rs_u, rc_u, al_u, be_u, ga_u);
    rs_u = uni(rs_u); rc_u = uni(rc_u); al_u = uni(al_u); be_u = uni(be_u); ga_u = uni(ga_u);
    // The phase at a segment's start is the one the previous segment's
    // rotations reached at its end (the same node); it is evaluated directly
    // only every ANCHOR segments (drift of a rotation chain: a few ulp per
    // step, ~1e-14 rad over ANCHOR segments, below the ~1e-12 rad rounding of
    // omega x + p0 itself at x ~ 1e3).
    constexpr int ANCHOR = 16;
    int since_anchor = ANCHOR;
    const bool pow2_div = ((neff - 1) & (neff - 2)) == 0;
    const double inv_div = 1.0 / (double)(neff - 1);
    double sg_prev_end = 0.0;
    double s_c = 0.0, c_c = 1.0;                  // phase carried to the next segment's start
    for (long sg = s_lo; sg < s_hi; ++sg) {
        const double a = __dadd_rn(__dmul_rn((double)sg, bstep), r_m);
        const double b = (sg + 1 == nseg) ? r_max : __dadd_rn(__dmul_rn((double)(sg + 1), bstep), r_m);
        // (b - a) / (neff - 1); a power-of-two divisor (neff = 5, 9) is an exact scaling
        const double step = pow2_div ? (b - a) * inv_div : (b - a) / (neff - 1);
        const double hseg = b - a;
        if (hseg < 1e-15) continue;               // zero width: its end phase is its start phase
        // a segment's start is the previous one's end node (the same expression), whose reciprocal is pending
        double ix0 = (sg > s_lo && sg_prev_end == a) ? ix_pend : (a > 1e-6 ? recip(a) : 0.0);
        sg_prev_end = b;
        if (ix0 != ix_pend && w_pend != 0.0) {    // the pending node is not this segment's first
            sweep2<LE>(acc, w_pend, ix_pend, 0.0, 0.0);
            w_pend = 0.0;
        }
        const bool small = fabs(om * hseg) < 1e-3;
        const double cmid = small ? cos(om * (a + b) / 2 + p0) : 0.0;
        // rotation by theta = omega * step and the Filon coefficients at theta
        double rs = 0.0, rc = 1.0, s0 = 0.0, c0 = 0.0;
        double al = 0.0, be = 0.0, ga = 0.0;
        if (!small) {
            if (sg + 1 == nseg) {
                filon_coeffs(om * step, rs, rc, al, be, ga);
            } else {
                rs = rs_u; rc = rc_u; al = al_u; be = be_u; ga = ga_u;
            }
            if (since_anchor >= ANCHOR) {
                sincos(om * a + p0, &s0, &c0);
                since_anchor = 0;
            } else {
                s0 = s_c;
                c0 = c_c;
            }
            ++since_anchor;
        } else {
            since_anchor = ANCHOR;                // no rotation chain through a non-oscillatory segment
        }
        double x0 = a;
        const auto inner = [&](int k) { return __dadd_rn(__dmul_rn((double)k, step), a); };
        const auto node = [&](int k) { return (k == neff - 1) ? b : inner(k); };
        const auto inv = [](double x) { return x > 1e-6 ? recip(x) : 0.0; };
        // one triple (x0 x1 x2) per pass, x2 left pending, its nodes x0, x1 swept
        // in the block form; in a segment that starts past the envelope cut-off
        // (a > 1e-6, so every node is) 1/x1 and 1/x2 come from one reciprocal
        // of x1 x2 (~2 ulp instead of ~1, at half the reciprocals)
        const auto triples = [&](auto live) {
#pragma unroll 1
            for (int t = 0; t < nt; ++t) {
                const double x1 = inner(2 * t + 1), x2 = node(2 * t + 2);
                double ix1, ix2;
                if constexpr (decltype(live)::value) {
                    const double r12 = recip(x1 * x2);
                    ix1 = x2 * r12;
                    ix2 = x1 * r12;
                } else {
                    ix1 = inv(x1);
                    ix2 = inv(x2);
                }
                const double s1 = fma(s0, rc, c0 * rs), c1 = fma(c0, rc, -(s0 * rs));
                const double s2 = fma(s1, rc, c1 * rs), c2 = fma(c1, rc, -(s1 * rs));
                double wa, wm, wb;
                triple_weights(x0, x1, x2, s0, c0, c1, s2, c2, small, cmid, om, p0, inv_om, al, be, ga, wa, wm,
                               wb);
                sweep2b<LE>(acc, w_pend + wa, ix0, wm, ix1);
                w_pend = wb;
                ix_pend = ix2;
                x0 = x2;
                ix0 = ix2;
                s0 = s2;
                c0 = c2;
            }
        };
        if (a > 1e-6)
            triples(std::true_type{});
        else
            triples(std::false_type{});
        s_c = s0;
        c_c = c0;
    }
    if (w_pend != 0.0) sweep2<LE>(acc, w_pend, ix_pend, 0.0, 0.0);
}

// Phases and the Filon / Levin choice of compute_outer_integral_oscillatory
// for one (job, sign).
struct TailSetup {
    Phase P;
    double om;
    int kind;     // 0: nothing to do (zeros), 1: Filon, 2: Levin or non-oscillatory
};

__device__ inline TailSetup tail_setup(const dwba_tail_job& J, int sgn, int NL) {
    TailSetup T{};
    const int Lm = J.L_max < NL - 1 ? J.L_max : NL - 1;
    bool any = false;
    for (int L = 1; L <= Lm; ++L) any |= (J.moments && fabs(J.moments[L]) > 1e-12);
    if (J.kind != DWBA_TAIL_OUTER || !any || J.r_max <= J.r_m + 1e-10) {
        T.kind = 0;
        return T;
    }
    Phase& P = T.P;
    P.ka = J.k_a; P.kb = J.k_b; P.ea = J.eta_a; P.eb = J.eta_b;
    P.la = (double)J.l_a; P.lb = (double)J.l_b;
    P.use_ea = fabs(J.eta_a) > 1e-15; P.use_eb = fabs(J.eta_b) > 1e-15;
    P.use_la = J.l_a > 0; P.use_lb = J.l_b > 0;
    P.kmin = J.k_a - J.k_b; P.kplus = J.k_a + J.k_b;
    const double pa = -J.l_a * PI / 2 + J.sigma_a + J.delta_a;
    const double pb = -J.l_b * PI / 2 + J.sigma_b + J.delta_b;
    P.cmin = pa - pb; P.cplus = pa + pb;
    const double mid = (J.r_m + J.r_max) / 2;
    T.om = fabs(dphi(P, sgn, mid));
    bool filon = false;
    if (!(T.om < 1e-10)) {
        const double h = (PI / 4) / T.om;
        filon = fabs(d2phi(P, sgn, mid)) * h * h < 0.1;
    }
    T.kind = filon ? 1 : 2;
    return T;
}

// ---------------------------------------------------------------------------
// Filon signs, all L: the phase functions do not depend on L (only the
// envelope M_L / r^(L+1) does), so a wavefront evaluates nodes, exponentials
// and Filon coefficients once and sweeps all L.  Only Re J is needed
// (dwba_outer_integral_1d keeps the real parts).  One wavefront per (pair,
// sign, radial part); part p sweeps multipoles L < NL >> p over its segments
// (see part_begin) and writes its row of partial sums, zeros for the L it
// does not sweep and for signs that are not Filon (the Levin kernel then
// writes part 0's row of its signs).
template <int NL>
__global__ __launch_bounds__(64, NL <= 32 ? 3 : 2) void tail_filon_kernel(const dwba_tail_job* __restrict__ jobs, TailParams TP,
                                                        double* __restrict__ jre) {
    __shared__ double s_acc[64][SUM_COLS + 1];
    // blockIdx.y = 0: the Phi_+ sign (k_a + k_b, usually many more segments) is dispatched first
    const int job = blockIdx.x, row = 1 - (int)blockIdx.y, sgn = row == 0 ? -1 : 1;
    const dwba_tail_job J = jobs[job];
    const int lane = threadIdx.x;
    const int Lm = J.L_max < NL - 1 ? J.L_max : NL - 1;
    double* out = jre + ((size_t)job * 2 + row) * NL;
    const TailSetup T = tail_setup(J, sgn, NL);
    if (T.kind == 0)
        for (int L = lane; L < NL; L += 64) out[L] = 0.0;
    if (T.kind != 1) return;
    const double om = T.om;
    const double r_m = J.r_m, r_max = J.r_max;
    const double delta_phi = PI / 4;
    double acc[NL];
#pragma unroll
    for (int L = 0; L < NL; ++L) acc[L] = 0.0;
    const double p0 = phi(T.P, sgn, r_m) - om * r_m;
    const double dr_seg = delta_phi / fabs(om);
    const double nraw_d = ceil((r_max - r_m) / dr_seg);
    const long nraw = nraw_d < 1.0 ? 1 : (long)nraw_d;
    const long nseg = nraw < TP.max_filon_segments ? nraw : TP.max_filon_segments;
    const double eff = fabs(om) * (r_max - r_m) / (double)(nseg > 1 ? nseg : 1);
    int neff = 5;
    if (eff > TP.max_effective_dphi) {
        const int scale = (int)ceil(eff / TP.max_effective_dphi);
        neff = 5 + 2 * (scale - 1);
        if (neff < 5) neff = 5;
        if (neff > 15) neff = 15;
    }
    const double bstep = (r_max - r_m) / (double)nseg;
    filon_sweep<NL>(neff, om, p0, r_m, r_max, bstep, nseg, 0, nseg, lane, acc);
    lane_sums_out<NL, NL>(acc, s_acc, lane, J.moments, Lm, out);
}

// Levin signs, all L (levin_segment's segments of compute_outer_integral_oscillatory:
// n_seg = ceil((r_max - r_m) / h), h = (pi/4) / |Phi'(mid)|, 8 Chebyshev-Lobatto
// nodes each): per segment one LU gives the node weights Re g (levin_weights),
// swept over all L like the Filon weights; lanes take segments round-robin.
// Non-oscillatory signs (|Phi'(mid)| < 1e-10, scipy quad in the reference)
// stay per L with the Gauss-Legendre panels.  Writes part 0's row.
template <int NL>
__global__ __launch_bounds__(64) void tail_levin_kernel(const dwba_tail_job* __restrict__ jobs, TailParams TP,
                                                        double* __restrict__ jre) {
    __shared__ double s_acc[64][SUM_COLS + 1];
    const int job = blockIdx.x, row = 1 - (int)blockIdx.y, sgn = row == 0 ? -1 : 1;
    const dwba_tail_job J = jobs[job];
    const int lane = threadIdx.x;
    const TailSetup T = tail_setup(J, sgn, NL);
    if (T.kind != 2) return;
    const int Lm = J.L_max < NL - 1 ? J.L_max : NL - 1;
    double* out = jre + ((size_t)job * 2 + row) * NL;
    const double r_m = J.r_m, r_max = J.r_max;
    if (T.om < 1e-10) {
        for (int L = 1; L <= Lm; ++L) {
            const double M = J.moments[L];
            double v = 0.0;
            if (fabs(M) > 1e-12) v = gauss_quad(Env{M, L + 1}, T.P, sgn, r_m, r_max, false, 0.0, lane).re;
            if (lane == 0) out[L] = v;
        }
        if (lane == 0) out[0] = 0.0;
        for (int L = Lm + 1 + lane; L < NL; L += 64) out[L] = 0.0;
        return;
    }
    const double h = (PI / 4) / T.om;
    const double nseg_d = ceil((r_max - r_m) / h);
    const long nseg = nseg_d < 1.0 ? 1 : (long)nseg_d;
    const double bstep = (r_max - r_m) / (double)nseg;
    double acc[NL];
#pragma unroll
    for (int L = 0; L < NL; ++L) acc[L] = 0.0;
    for (long s = lane; s < nseg; s += 64) {
        const double a = __dadd_rn(__dmul_rn((double)s, bstep), r_m);
        const double b = (s + 1 == nseg) ? r_max : __dadd_rn(__dmul_rn((double)(s + 1), bstep), r_m);
        double gre[8], ix[8];
        if (!levin_weights(T.P, sgn, a, b, TP.lev, gre, ix)) continue;
#pragma unroll
        for (int k = 0; k < 8; k += 2) sweep2<NL>(acc, gre[k], ix[k], gre[k + 1], ix[k + 1]);
    }
    lane_sums_out<NL, NL>(acc, s_acc, lane, J.moments, Lm, out);
}

// tail[L] = 1/2 (Re J_- - Re J_+) + analytic tail at r_max  (or the legacy analytic tail)
__global__ void tail_combine_kernel(const dwba_tail_job* __restrict__ jobs, int n_jobs, int NL,
                                    const double* __restrict__ jre) {
    const int job = blockIdx.x;
    const dwba_tail_job J = jobs[job];
    for (int L = threadIdx.x; L <= J.L_max; L += blockDim.x) {
        double res = 0.0;
        const double M = J.moments ? J.moments[L] : 0.0;
        if (J.kind != DWBA_TAIL_NONE && L >= 1 && fabs(M) > 1e-12) {
            if (J.kind == DWBA_TAIL_LEGACY) {
                res = analytic_tail(J.r_m, J, L, M);
            } else {
                double outer = 0.0;
                if (!(J.r_max <= J.r_m + 1e-10) && L < NL) {
                    const double jm = jre[((size_t)job * 2 + 0) * NL + L];
                    const double jp = jre[((size_t)job * 2 + 1) * NL + L];
                    outer = 0.5 * (jm - jp);
                }
                res = outer + analytic_tail(J.r_max, J, L, M);
            }
        }
        J.out[L] = res;
    }
}

int tail_nl_bucket(int L_max) {
    // 31: L_max = 30 (the common multipole range) sweeps no unused L = 31
    static const int buckets[] = {2, 4, 8, 16, 24, 31, 32, 48, 64};
    for (int b : buckets)
        if (L_max + 1 <= b) return b;
    return -1;
}

}  // namespace
}  // namespace dwba

extern "C" size_t dwba_outer_tail_workspace_bytes(int32_t n_jobs, int32_t L_max) {
    const int nl = dwba::tail_nl_bucket(L_max);
    if (n_jobs <= 0 || nl < 0) return 0;
    return (size_t)n_jobs * 2 * nl * sizeof(double);
}

extern "C" int32_t dwba_outer_tail_f64(int32_t n_jobs, const dwba_tail_job* d_jobs, int32_t L_max,
                                       int32_t max_filon_segments, double max_effective_dphi,
                                       const double* h_levin_D, const double* h_levin_x, void* d_workspace,
                                       size_t workspace_bytes, void* stream) {
    DWBA_REQUIRE(dwba::initialised(), "dwba_outer_tail_f64: call dwba_init first");
    DWBA_REQUIRE(n_jobs > 0 && d_jobs && L_max >= 0 && L_max <= 63 && h_levin_D && h_levin_x,
                 "dwba_outer_tail_f64: bad arguments");
    DWBA_REQUIRE(max_filon_segments >= 1 && max_effective_dphi > 0, "dwba_outer_tail_f64: bad Filon limits");
    const int nl = dwba::tail_nl_bucket(L_max);
    const size_t need = (size_t)n_jobs * 2 * nl * sizeof(double);
    if (!d_workspace || workspace_bytes < need)
        return dwba::set_error(DWBA_ERR_WORKSPACE, "dwba_outer_tail_f64: workspace %zu < %zu bytes",
                               workspace_bytes, need);
    dwba::TailParams tp;
    tp.max_filon_segments = max_filon_segments;
    tp.max_effective_dphi = max_effective_dphi;
    for (int i = 0; i < 8; ++i) {
        tp.lev.xr[i] = h_levin_x[i];
        for (int j = 0; j < 8; ++j) tp.lev.D[i][j] = h_levin_D[i * 8 + j];
    }
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    double* jre = reinterpret_cast<double*>(d_workspace);
    switch (nl) {
#define C(n) case n: hipLaunchKernelGGL(dwba::tail_filon_kernel<n>, dim3(n_jobs, 2), dim3(64), 0, s, d_jobs, tp, jre); break;
        C(2) C(4) C(8) C(16) C(24) C(31) C(32) C(48) C(64)
#undef C
        default: return dwba::set_error(DWBA_ERR_ARG, "dwba_outer_tail_f64: L_max %d", L_max);
    }
    DWBA_LAUNCH_CHECK("tail_filon_kernel");
    switch (nl) {
#define C(n) case n: hipLaunchKernelGGL(dwba::tail_levin_kernel<n>, dim3(n_jobs, 2), dim3(64), 0, s, d_jobs, tp, jre); break;
        C(2) C(4) C(8) C(16) C(24) C(31) C(32) C(48) C(64)
#undef C
        default: break;
    }
    DWBA_LAUNCH_CHECK("tail_levin_kernel");
    hipLaunchKernelGGL(dwba::tail_combine_kernel, dim3(n_jobs), dim3(64), 0, s, d_jobs, n_jobs, nl, jre);
    DWBA_LAUNCH_CHECK("tail_combine_kernel");
    return DWBA_OK;
}
