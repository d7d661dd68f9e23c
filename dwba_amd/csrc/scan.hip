// Whole-scan engine of the C ABI: the per-pair host decisions of
// radial_ME_all_L for every channel pair of a scan, made once, and a fixed
// launch sequence over the library's kernels.  All device data of a scan
// (grids, potentials, waves, alpha/beta vectors, tables, results) lives in
// one allocation made by dwba_scan_plan; dwba_scan_launch allocates nothing.
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <new>
#include <numeric>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "common.hpp"

namespace dwba {
namespace scan_detail {

constexpr int LD_PAD = 32;          // vectors of an energy are padded to a multiple of 32 points
constexpr int GROUP_SLICE = 4;      // members per workgroup of the fused Filon pass
constexpr int COUL_NODES = 10002;   // Coulomb phase-node buffer (generate_phase_nodes caps at 10001)
constexpr int LANE_LB = 8;          // multipole block of the split lane contraction (L_max < 4)
constexpr int LANE_WAVES = 2;       // its occupancy (waves per SIMD)
constexpr int STAGE_LB = 4;         // multipole block of the LDS-staged lane contraction (L_max >= 4)
constexpr double PI = 3.141592653589793;

// ------------------------------------------------------------------ tables
// Clenshaw-Curtis reference nodes / weights on [-1, 1], nodes from +1 down to
// -1.  n == 5: the module table of oscillatory_integrals.py:120-142 (last
// cosine term with weight 1/2, weights summing to 31/15); otherwise
// clenshaw_curtis_nodes(n, -1, 1) (oscillatory_integrals.py:1539-1618).
bool cc_reference(int n, std::vector<double>& x, std::vector<double>& w) {
    if (n < 2 || n > DWBA_MAX_CC) return false;
    x.assign(n, 0.0);
    w.assign(n, 0.0);
    std::vector<double> th(n);
    for (int j = 0; j < n; ++j) th[j] = PI * j / (n - 1);
    if (n == 5) {
        const int jm = (n - 1) / 2;
        for (int j = 0; j < n; ++j) {
            x[j] = std::cos(th[j]);
            double ws = 0.0;
            for (int v = 1; v <= jm; ++v) ws += std::cos(2 * v * th[j]) / (4 * v * v - 1);
            ws += 0.5 * std::cos(2 * jm * th[j]) / (4 * jm * jm - 1);
            w[j] = (2.0 / (n - 1)) * (1 - 2 * ws);
        }
        return true;
    }
    if (n == 2) {
        x = {1.0, -1.0};
        w = {1.0, 1.0};
        return true;
    }
    const int m = n - 1, jm = m / 2;
    for (int j = 0; j < n; ++j) {
        x[j] = 0.5 * 2.0 * (std::cos(th[j]) + 1) + -1.0;
        double s = 0.0;
        for (int v = 1; v <= jm; ++v) s += (2.0 / (1 - 4 * v * v)) * std::cos(2 * v * th[j]);
        double wj = 1.0 + s;
        if (m % 2 == 0) wj += std::cos(m * th[j]) / (1 - m * m);
        w[j] = wj * (2.0 / m);
    }
    w[0] /= 2;
    w[n - 1] /= 2;
    return true;
}

// Chebyshev differentiation matrix on the descending Lobatto points
// cos(pi i / 7) and the ascending node positions of the 8-node Levin rule
// (oscillatory_integrals.py:970-1017, :1145-1147).
void levin_tables(double D[64], double xa[8]) {
    const int n = 8;
    double x[8], c[8];
    for (int i = 0; i < n; ++i) {
        x[i] = std::cos(PI * i / (n - 1));
        c[i] = (i == 0 || i == n - 1) ? 2.0 : 1.0;
    }
    for (int i = 0; i < n; ++i) {
        double row[8];
        for (int j = 0; j < n; ++j) {
            if (i == j) { row[j] = 0.0; continue; }
            const double sg = ((i + j) & 1) ? -1.0 : 1.0;
            row[j] = (c[i] / c[j]) * sg / (x[i] - x[j]);
        }
        // the diagonal is minus the row sum, added in numpy's pairwise order for 8 terms
        const double s = ((row[0] + row[1]) + (row[2] + row[3])) + ((row[4] + row[5]) + (row[6] + row[7]));
        row[i] = -s;
        for (int j = 0; j < n; ++j) D[i * n + j] = row[j];
    }
    for (int i = 0; i < n; ++i) xa[i] = x[n - 1 - i];
}

// DWBA_MAX_FILON_SEGMENTS (>= 128, default 4096) and DWBA_MAX_EFFECTIVE_DPHI
// (>= pi/16, default pi/2), read as oscillatory_integrals.py:54-80 does.
void filon_limits(int& seg, double& dphi) {
    seg = 4096;
    dphi = PI / 2;
    if (const char* s = std::getenv("DWBA_MAX_FILON_SEGMENTS")) {
        char* end = nullptr;
        const long v = std::strtol(s, &end, 10);
        if (end != s) {
            while (*end == ' ' || *end == '\t' || *end == '\n') ++end;
            if (*end == '\0') seg = (int)std::max(128L, std::min(v, 1L << 30));
        }
    }
    if (const char* s = std::getenv("DWBA_MAX_EFFECTIVE_DPHI")) {
        char* end = nullptr;
        const double v = std::strtod(s, &end);
        if (end != s) {
            while (*end == ' ' || *end == '\t' || *end == '\n') ++end;
            if (*end == '\0' && std::isfinite(v)) dphi = std::max(PI / 16, v);
        }
    }
}

// ------------------------------------------------------------------ grid statistics
// Prefix statistics of one grid, so that every per-pair decision below is
// O(log N) (the Coulomb phase-step maxima are built once per (k, eta)).
struct GridStats {
    const double* r = nullptr;
    const double* w_trapz = nullptr;
    int n = 0;
    bool pos_inc = false;
    std::vector<double> dr_max;     // max |dr| over steps 0..k
    std::vector<double> dl_mean, dl_max, dl_min;
    std::map<std::pair<double, double>, std::vector<double>> coul;

    void init(const double* r_, const double* wt, int n_) {
        r = r_;
        w_trapz = wt;
        n = n_;
        pos_inc = n >= 1 && r[0] > 0.0;
        for (int i = 0; i + 1 < n && pos_inc; ++i) pos_inc = r[i + 1] - r[i] > 0.0;
        dr_max.resize(n > 1 ? n - 1 : 0);
        double m = 0.0;
        for (int i = 0; i + 1 < n; ++i) {
            m = i ? std::max(m, std::fabs(r[i + 1] - r[i])) : std::fabs(r[i + 1] - r[i]);
            dr_max[i] = m;
        }
        if (pos_inc && n >= 2) {
            dl_mean.resize(n - 1);
            dl_max.resize(n - 1);
            dl_min.resize(n - 1);
            double s = 0.0, mx = 0.0, mn = 0.0, lp = std::log(r[0]);
            for (int i = 0; i + 1 < n; ++i) {
                const double ln = std::log(r[i + 1]);
                const double d = ln - lp;
                lp = ln;
                s += d;
                mx = i ? std::max(mx, d) : d;
                mn = i ? std::min(mn, d) : d;
                dl_mean[i] = s / (double)(i + 1);
                dl_max[i] = mx;
                dl_min[i] = mn;
            }
        }
    }

    // _is_log_spaced_grid on r[:m] (oscillatory_integrals.py:237-255): every
    // log step within rel_tol of the mean step.
    bool log_spaced(int m, double rel_tol = 5e-3) const {
        if (m < 3 || !pos_inc) return false;
        const int k = m - 2;
        const double mean = dl_mean[k];
        if (!std::isfinite(mean) || mean <= 0.0) return false;
        const double dev = std::max(dl_max[k] - mean, mean - dl_min[k]);
        return dev <= rel_tol * mean;
    }

    // Largest phase change per step over r[:m] (check_phase_sampling,
    // oscillatory_integrals.py:1456-1509).
    double max_phase_step(int m, double k_total, double eta_total) {
        if (m < 2) return 0.0;
        if (std::fabs(eta_total) > 1e-10) {
            auto key = std::make_pair(k_total, eta_total);
            auto it = coul.find(key);
            if (it == coul.end()) {
                std::vector<double> pm(n - 1);
                double mx = 0.0;
                for (int i = 0; i + 1 < n; ++i) {
                    const double rm = std::max(0.5 * (r[i] + r[i + 1]), 1e-10);
                    const double v = std::fabs((k_total + eta_total / rm) * (r[i + 1] - r[i]));
                    mx = i ? std::max(mx, v) : v;
                    pm[i] = mx;
                }
                it = coul.emplace(key, std::move(pm)).first;
            }
            return it->second[m - 2];
        }
        return std::fabs(k_total) * dr_max[m - 2];
    }
};

// Last index < n where a is not finite, -1 if none.
int last_nonfinite(const double* a, int n) {
    for (int i = n - 1; i >= 0; --i)
        if (!std::isfinite(a[i])) return i;
    return -1;
}

// First index i with a[i] >= v (numpy searchsorted, side "left").
int search_left(const double* a, int n, double v) { return (int)(std::lower_bound(a, a + n, v) - a); }

// Index of the 99% cumulative probability of u^2 (trapezoid weights), 0
// when the density vanishes (dwba_matrix_elements.py:962-985).
int extent99(const double* u, const double* wt, int n) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += u[i] * u[i];
    if (!(s > 1e-30)) return 0;
    std::vector<double> pc(n);
    double acc = 0.0;
    for (int i = 0; i < n; ++i) {
        acc += (u[i] * u[i]) * wt[i];
        pc[i] = acc;
    }
    const double den = pc[n - 1] + 1e-300;
    for (int i = 0; i < n; ++i) pc[i] /= den;
    return search_left(pc.data(), n, 0.99);
}

// idx_limit of radial_ME_all_L (dwba_matrix_elements.py:909-1022): the
// earlier match point, raised past the centrifugal turning point (at least
// min_grid_fraction of the grid), extended to the 99% extent of u_i and u_f,
// then pushed out (at most 500 points) until
// max(|U_i|, |U_f|) + l(l+1)/(2 r^2) < 0.03 k_min^2 / 2.
int split_index(const GridStats& gs, const double* U_i, const double* U_f, int ext, double k_i, double k_f, int l_i,
                int l_f, int im_i, int im_f, double mgf) {
    const double* r = gs.r;
    const int n = gs.n;
    int idx = n;
    if (0 < im_i && im_i < n) idx = std::min(idx, im_i + 1);
    if (0 < im_f && im_f < n) idx = std::min(idx, im_f + 1);
    const double kmin = std::min(k_i, k_f);
    const double k_eff = kmin > 0.1 ? kmin : std::max(k_i, k_f);
    const int lw = std::max(l_i, l_f);
    int floor_idx;
    if (k_eff > 1e-3) {
        const int it = std::min(search_left(r, n, (lw + 0.5) / k_eff), n - 20);
        floor_idx = std::min(std::max(it + 20, (int)(n * mgf)), n);
    } else {
        floor_idx = std::max(1, (int)(n * mgf));
    }
    idx = std::max(idx, floor_idx);
    if (ext > idx) idx = ext + 1;
    const double kin = 0.5 * (kmin * kmin);
    if (kin > 1e-10) {
        const double cent = (double)(lw * (lw + 1));
        auto veff = [&](int t) {
            const double ui = std::fabs(U_i[t]);
            const double uf = U_f ? std::fabs(U_f[t]) : 0.0;
            return std::max(ui, uf) + cent / (2.0 * (r[t] * r[t]));
        };
        if (veff(std::max(0, idx - 1)) / kin > 0.03) {
            const int t1 = std::min(idx + 500, n);
            for (int t = idx; t < t1; ++t) {
                if (veff(t) / kin < 0.03) {
                    idx = t + 1;
                    break;
                }
            }
        }
    }
    return idx;
}

// _adaptive_phase_increment_for_grid (oscillatory_integrals.py:1716-1755).
double phase_increment(GridStats& gs, int n1, double k_total, double requested, double eta_total) {
    const double threshold = PI / 4, min_inc = PI / 16;
    const double d_req = std::min(std::max(min_inc, requested), threshold);
    if (k_total < 1e-10 || n1 < 2) return d_req;
    const double mp = gs.max_phase_step(n1, k_total, eta_total);
    if (!(mp > threshold)) return d_req;
    const double scale = threshold / std::max(mp, 1e-12);
    const double v = d_req * std::max(0.25, std::min(1.0, scale));
    return std::min(std::max(v, min_inc), threshold);
}

struct Layout {
    int mode = DWBA_PREP_PLAIN;
    double p_start = 0.0, p_step = 0.0, p_end = 0.0;
    int n_pnodes = 0;
    bool coulomb = false;
    double k_total = 0.0, eta_total = 0.0;
};

// Branch structure of oscillatory_kernel_integral_2d's filon /
// filon_exchange methods (oscillatory_integrals.py:2441-2628).
Layout filon_layout(GridStats& gs, int n1, double k_total, double phase_inc, double eta_total) {
    Layout L;
    if (n1 < 2) {
        L.mode = DWBA_PREP_ZERO;
        return L;
    }
    const double dphi = phase_increment(gs, n1, k_total, phase_inc, eta_total);
    if (k_total < 1e-6) {
        L.mode = DWBA_PREP_STANDARD;
        return L;
    }
    const double a = gs.r[0], b = gs.r[n1 - 1];
    if (std::fabs(eta_total) < 1e-10) {
        const double dr = dphi / k_total;
        const double q = std::ceil((b - a) / dr);
        if (!(q >= 1.0) || q > 2e9) {
            L.mode = DWBA_PREP_STANDARD;
            return L;
        }
        const int n_pts = (int)q + 1;
        const double step = (b - a) / (n_pts - 1);
        if (!(step > 1e-12)) {
            L.mode = DWBA_PREP_STANDARD;
            return L;
        }
        L.mode = DWBA_PREP_FILON;
        L.p_start = a;
        L.p_step = step;
        L.p_end = b;
        L.n_pnodes = n_pts;
        L.k_total = k_total;
        L.eta_total = eta_total;
        return L;
    }
    if (!((b - a) > 1e-12)) {
        L.mode = DWBA_PREP_STANDARD;
        return L;
    }
    L.mode = DWBA_PREP_FILON;
    L.p_start = a;
    L.p_step = dphi;
    L.p_end = b;
    L.coulomb = true;
    L.k_total = k_total;
    L.eta_total = eta_total;
    return L;
}

// Nodes one pair's outer tail sweeps (Filon + Levin), to start the longest
// tails first (compute_outer_integral_oscillatory, oscillatory_integrals.py:1361-1451).
long tail_sweep_nodes(double k_i, int l_i, double eta_i, double k_f, int l_f, double eta_f, double r_m, double r_max,
                      int max_seg, double max_dphi) {
    if (r_max <= r_m + 1e-10) return 0;
    long total = 0;
    const double mid = 0.5 * (r_m + r_max);
    for (int s = 0; s < 2; ++s) {
        const double sgn = s ? 1.0 : -1.0;
        double d1 = sgn < 0 ? (k_i - k_f) : (k_i + k_f), d2 = 0.0;
        if (std::fabs(eta_i) > 1e-15) { d1 += eta_i / mid; d2 -= eta_i / (mid * mid); }
        if (std::fabs(eta_f) > 1e-15) { d1 += sgn * eta_f / mid; d2 -= sgn * eta_f / (mid * mid); }
        if (l_i > 0) { d1 += l_i * (l_i + 1) / (2.0 * k_i * mid * mid); d2 -= l_i * (l_i + 1) / (k_i * mid * mid * mid); }
        if (l_f > 0) {
            d1 += sgn * l_f * (l_f + 1) / (2.0 * k_f * mid * mid);
            d2 -= sgn * l_f * (l_f + 1) / (k_f * mid * mid * mid);
        }
        const double om = std::fabs(d1);
        if (om < 1e-10) continue;
        const double h = (PI / 4) / om, span = r_max - r_m;
        const double segs = std::min(std::max(1.0, std::ceil(span / h)), 1e9);
        if (std::fabs(d2) * h * h < 0.1) {
            const long nseg = std::min((long)segs, (long)max_seg);
            const double eff = om * span / nseg;
            int neff = 5;
            if (eff > max_dphi) neff = std::min(15, std::max(5, 5 + 2 * ((int)std::ceil(eff / max_dphi) - 1)));
            total += nseg * (neff - 1) + 1;
        } else {
            total += 8 * (long)segs;
        }
    }
    return total;
}

// The host decisions for one pair (plan_pair of the drop-in).
struct Plan {
    int energy = 0;
    int N = 0, N1 = 0, mode = DWBA_PREP_PLAIN, tail = DWBA_TAIL_NONE, rows = 0, cols = 0, log1 = 0, log2 = 0;
    bool osc = false;
    Layout lay;
    double k_i = 0, k_f = 0, delta_i = 0, delta_f = 0, eta_i = 0, eta_f = 0, sigma_i = 0, sigma_f = 0;
    int l_i = 0, l_f = 0;
    double r_match = 0, r_max = 0;
};

// Bump allocator over one device allocation (offsets first, base later).
// The first grid takes offset 0, so 0 marks "none" for optional buffers.
struct Arena {
    size_t top = 0;
    size_t take(size_t bytes) {
        const size_t off = top;
        top += align_up(bytes == 0 ? 1 : bytes, 256);
        return off;
    }
};

}  // namespace scan_detail
}  // namespace dwba

// One planned scan: its device allocation, launch tables and launch parameters.
struct dwba_scan {
    int device = 0;
    int32_t n_pairs = 0, L_max = 0;
    void* base = nullptr;
    size_t bytes = 0;
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // prep
    dwba_prep_job* prep = nullptr;
    const double* coul = nullptr;
    int any_coul = 0;
    std::vector<double> cc_x, cc_w;
    // grouped Filon
    const dwba_filon_group* groups = nullptr;
    const dwba_filon_group* slices = nullptr;
    const int32_t* members = nullptr;
    int32_t n_groups = 0, n_slices = 0;
    // moments and tails
    int32_t n_sets = 0, n_tail = 0;
    const double* const* mom_ptr[4] = {nullptr, nullptr, nullptr, nullptr};
    const int32_t* mom_n = nullptr;
    double* moments = nullptr;
    const dwba_tail_job* tail_jobs = nullptr;
    void* tail_ws = nullptr;
    size_t tail_ws_bytes = 0;
    int max_seg = 4096;
    double max_dphi = 1.5707963267948966;
    double levin_D[64], levin_x[8];
    // contraction
    int32_t n_jobs = 0;
    int lane_lb = 8;                   // 4: the LDS-staged kernel on jobs padded to whole groups
    bool stage = false;
    const dwba_plain_job* jobs = nullptr;
    const double* const* tabs = nullptr;
    const int32_t* out_index = nullptr;
    int32_t n_gemv = 0, n_dots = 0, n_items16 = 0;
    const dwba_dgemv_energy* gemv = nullptr;
    const dwba_ddot_pair* dots = nullptr;
    const int32_t* items16 = nullptr;
    // finalize and results
    const dwba_finalize_pair* fin = nullptr;
    double* out = nullptr;
    int64_t info[8] = {0, 0, 0, 0, 0, 0, 0, 0};
};

namespace {

using namespace dwba;
using namespace dwba::scan_detail;

void scan_free(dwba_scan* s) {
    if (!s) return;
    if (s->exec) (void)hipGraphExecDestroy(s->exec);
    if (s->graph) (void)hipGraphDestroy(s->graph);
    if (s->ev_fork) (void)hipEventDestroy(s->ev_fork);
    if (s->ev_join) (void)hipEventDestroy(s->ev_join);
    if (s->side) (void)hipStreamDestroy(s->side);
    if (s->base) (void)hipFree(s->base);
    delete s;
}

template <class T>
T* at(void* base, size_t off) {
    return reinterpret_cast<T*>(static_cast<char*>(base) + off);
}

int enqueue(dwba_scan* s, hipStream_t st) {
    int rc;
    if (s->n_tail) {
        // the tail chain reads only bound densities, wave parameters and its
        // own moments: it runs on a side stream forked here and joined before
        // the finalize step
        DWBA_HIP_CHECK(hipEventRecord(s->ev_fork, st));
        DWBA_HIP_CHECK(hipStreamWaitEvent(s->side, s->ev_fork, 0));
        rc = dwba_moments_f64(s->n_sets, s->mom_ptr[0], s->mom_ptr[1], s->mom_ptr[2], s->mom_ptr[3], s->mom_n,
                              s->L_max, s->moments, s->side);
        if (rc) return rc;
        rc = dwba_outer_tail_f64(s->n_tail, s->tail_jobs, s->L_max, s->max_seg, s->max_dphi, s->levin_D, s->levin_x,
                                 s->tail_ws, s->tail_ws_bytes, s->side);
        if (rc) return rc;
        DWBA_HIP_CHECK(hipEventRecord(s->ev_join, s->side));
    }
    rc = dwba_fullsplit_prep_f64(s->n_pairs, s->prep, s->coul, s->any_coul, s->cc_x.data(), s->cc_w.data(),
                                 (int32_t)s->cc_x.size(), st);
    if (rc) return rc;
    if (s->n_groups) {
        rc = dwba_filon_cells_f64(s->n_groups, s->groups, s->prep, s->cc_x.data(), s->cc_w.data(),
                                  (int32_t)s->cc_x.size(), st);
        if (rc) return rc;
        rc = dwba_filon_fused_f64(s->n_slices, s->slices, s->prep, s->members, st);
        if (rc) return rc;
    }
    rc = s->stage ? dwba_il_semisep_lanes_stage_f64(s->n_jobs, s->jobs, s->tabs, s->out_index, s->L_max, STAGE_LB,
                                                    s->out, st)
                  : dwba_il_semisep_lanes_split_f64(s->n_jobs, s->jobs, s->tabs, s->out_index, s->L_max, LANE_LB,
                                                    LANE_WAVES, s->out, st);
    if (rc) return rc;
    if (s->n_gemv) {
        rc = dwba_semisep_gemv_f64(s->n_gemv, s->gemv, s->L_max, st);
        if (rc) return rc;
        rc = s->L_max <= 31 ? dwba_ddot_stage_f64(s->n_items16, s->items16, s->dots, s->L_max, s->out, st)
                            : dwba_ddot_f64(s->n_dots, s->dots, s->L_max, s->out, st);
        if (rc) return rc;
    }
    if (s->n_tail) DWBA_HIP_CHECK(hipStreamWaitEvent(st, s->ev_join, 0));
    return dwba_batch_finalize(s->n_pairs, s->fin, s->L_max, s->out, st);
}

// Validates a scan description and makes the per-pair decisions of
// radial_ME_all_L (plan_pair: dwba_matrix_elements.py:884-1108).
int make_plans(const char* who, int32_t n_energies, const dwba_scan_energy* E, int32_t n_waves,
               const dwba_scan_wave* W, int32_t n_pairs, const dwba_scan_pair* P, const dwba_scan_options* opt,
               std::vector<GridStats>& gs, std::vector<int>& wave_energy, std::vector<Plan>& plan) {
    const int Lm = opt->L_max;
    const bool use_osc = opt->use_oscillatory_quadrature != 0;
    DWBA_REQUIRE(Lm >= 0 && Lm <= 63, "%s: L_max %d outside [0, 63]", who, Lm);
    DWBA_REQUIRE(opt->method == DWBA_METHOD_LEGACY || opt->method == DWBA_METHOD_ADVANCED ||
                     opt->method == DWBA_METHOD_FULL_SPLIT,
                 "%s: unknown oscillatory method %d", who, opt->method);
    DWBA_REQUIRE(std::isfinite(opt->phase_increment) && std::isfinite(opt->min_grid_fraction) &&
                     std::isfinite(opt->k_threshold),
                 "%s: non-finite option", who);
    // energies: grid statistics and the 99% extent of u_i
    gs.assign(n_energies, GridStats());
    std::vector<int> ext_i(n_energies);
    for (int e = 0; e < n_energies; ++e) {
        const dwba_scan_energy& en = E[e];
        DWBA_REQUIRE(en.r && en.w_simpson && en.w_trapz && en.V_core && en.U_i && en.u_i && en.n >= 2,
                     "%s: energy %d: missing arrays or n < 2", who, e);
        gs[e].init(en.r, en.w_trapz, en.n);
        ext_i[e] = extent99(en.u_i, en.w_trapz, en.n);
    }
    wave_energy.assign(n_waves, -1);
    for (int k = 0; k < n_pairs; ++k) {
        const dwba_scan_pair& p = P[k];
        DWBA_REQUIRE(p.energy >= 0 && p.energy < n_energies && p.wave_i >= 0 && p.wave_i < n_waves &&
                         p.wave_f >= 0 && p.wave_f < n_waves && p.u_f,
                     "%s: pair %d: energy or wave index out of range, or u_f missing", who, k);
        const int N = E[p.energy].n;
        DWBA_REQUIRE(p.n_u_f >= N, "%s: pair %d: u_f has %d points but the grid has %d", who, k, p.n_u_f, N);
        for (int w : {p.wave_i, p.wave_f}) {
            DWBA_REQUIRE(W[w].chi && W[w].n >= N, "%s: wave %d: chi missing or shorter than the grid (%d < %d)", who,
                         w, W[w].n, N);
            DWBA_REQUIRE(wave_energy[w] < 0 || wave_energy[w] == p.energy,
                         "%s: wave %d is used by pairs of energies %d and %d", who, w, wave_energy[w], p.energy);
            wave_energy[w] = p.energy;
        }
    }
    plan.assign(n_pairs, Plan());
    std::map<std::pair<int, const double*>, int> ext_f;       // (energy, u_f) -> 99% extent
    for (int k = 0; k < n_pairs; ++k) {
        const dwba_scan_pair& p = P[k];
        const dwba_scan_energy& en = E[p.energy];
        const dwba_scan_wave& ci = W[p.wave_i];
        const dwba_scan_wave& cf = W[p.wave_f];
        const int N = en.n;
        auto ef = ext_f.find({p.energy, p.u_f});
        if (ef == ext_f.end()) ef = ext_f.emplace(std::make_pair(p.energy, p.u_f), extent99(p.u_f, en.w_trapz, N)).first;
        Plan& q = plan[k];
        q.energy = p.energy;
        q.N = N;
        q.k_i = ci.k;
        q.k_f = cf.k;
        q.l_i = ci.l;
        q.l_f = cf.l;
        q.N1 = split_index(gs[p.energy], en.U_i, en.U_f, std::max(ext_i[p.energy], ef->second), ci.k, cf.k, ci.l, cf.l,
                           ci.idx_match, cf.idx_match, opt->min_grid_fraction);
        q.eta_i = ci.eta;
        q.eta_f = cf.eta;
        const double k_total = ci.k + cf.k;
        q.osc = use_osc && k_total > opt->k_threshold;
        if (q.osc) q.lay = filon_layout(gs[p.energy], q.N1, k_total, opt->phase_increment, ci.eta + cf.eta);
        q.mode = q.lay.mode;
        const bool exc = p.excitation != 0;
        if (opt->method == DWBA_METHOD_LEGACY)
            q.tail = (use_osc && q.N1 < N - 10 && exc && Lm >= 1) ? DWBA_TAIL_LEGACY : DWBA_TAIL_NONE;
        else
            q.tail = (q.osc && q.N1 < N - 10 && exc && Lm >= 1) ? DWBA_TAIL_OUTER : DWBA_TAIL_NONE;
        if (q.mode == DWBA_PREP_PLAIN) {
            q.rows = q.N1;
            q.cols = q.N1;
        } else if (q.mode == DWBA_PREP_ZERO) {
            q.rows = 0;
            q.cols = N;
        } else {
            q.rows = q.N1;
            q.cols = N;
        }
        q.log1 = gs[p.energy].log_spaced(q.N1);
        q.log2 = gs[p.energy].log_spaced(N);
        q.delta_i = ci.phase_shift;
        q.delta_f = cf.phase_shift;
        q.sigma_i = ci.sigma_l;
        q.sigma_f = cf.sigma_l;
        q.r_match = q.N1 > 0 ? en.r[q.N1 - 1] : en.r[N - 1];
        q.r_max = en.r[N - 1];
    }
    return DWBA_OK;
}

// Plans the scan into s (allocation and uploads included); s is freed by the caller on error.
int plan_scan(dwba_scan* s, int32_t n_energies, const dwba_scan_energy* E, int32_t n_waves, const dwba_scan_wave* W,
              int32_t n_pairs, const dwba_scan_pair* P, const dwba_scan_options* opt) {
    const int Lm = opt->L_max, nl = Lm + 1;
    s->n_pairs = n_pairs;
    s->L_max = Lm;
    DWBA_REQUIRE(cc_reference(opt->CC_nodes, s->cc_x, s->cc_w), "dwba_scan_plan: CC_nodes %d outside [2, %d]",
                 opt->CC_nodes, DWBA_MAX_CC);
    // the LDS-staged lane kernel where it applies: groups of G jobs of one energy
    const int stage_group = dwba_semisep_lane_stage_group(Lm, STAGE_LB);
    s->stage = stage_group > 0;
    s->lane_lb = s->stage ? STAGE_LB : LANE_LB;
    const int stride = dwba_semisep_lane_table_stride(Lm, s->lane_lb);
    DWBA_REQUIRE(stride > 0, "dwba_scan_plan: L_max %d outside [0, 63]", Lm);
    const int n_slots = s->stage ? n_pairs + n_energies * (stage_group - 1) : n_pairs;   // job slots with padding
    std::vector<GridStats> gs;
    std::vector<int> wave_energy;
    std::vector<Plan> plan;
    int rc0 = make_plans("dwba_scan_plan", n_energies, E, n_waves, W, n_pairs, P, opt, gs, wave_energy, plan);
    if (rc0) return rc0;
    std::vector<int> ld(n_energies);
    int max_n = 0;
    for (int e = 0; e < n_energies; ++e) {
        ld[e] = (E[e].n + LD_PAD - 1) / LD_PAD * LD_PAD;
        max_n = std::max(max_n, E[e].n);
    }

    // ---- shared direct beta: energies whose pairs all share one u_f take the
    // direct family of their Filon / standard pairs from t_L = K_L beta (once
    // per energy) and a dot per pair
    std::vector<std::vector<int>> by_energy(n_energies);
    for (int k = 0; k < n_pairs; ++k) by_energy[P[k].energy].push_back(k);
    std::vector<char> shared_dir(n_pairs, 0), no_beta(n_pairs, 0);
    std::vector<int> gemv_energies, gemv_owner, gemv_rows;
    std::vector<std::vector<int>> gemv_pairs;
    for (int e = 0; e < n_energies; ++e) {
        std::vector<int> ks;
        bool one_uf = true;
        for (int k : by_energy[e]) {
            if (P[k].u_f != P[by_energy[e][0]].u_f) one_uf = false;
            if (plan[k].mode == DWBA_PREP_FILON || plan[k].mode == DWBA_PREP_STANDARD) ks.push_back(k);
        }
        if (ks.empty() || !one_uf) continue;
        int rows = 0;
        for (int k : ks) {
            rows = std::max(rows, plan[k].N1);
            shared_dir[k] = 1;
            no_beta[k] = k != ks[0];
        }
        gemv_energies.push_back(e);
        gemv_owner.push_back(ks[0]);
        gemv_rows.push_back(rows);
        gemv_pairs.push_back(ks);
    }

    // ---- Filon groups: pairs of one energy with the same split index, phase
    // layout and interpolation kinds share their cell sums
    std::map<std::tuple<int, int, int, int, double, double, double, int, int, double, double>, int> gkey;
    std::vector<std::vector<int>> groups;
    for (int k = 0; k < n_pairs; ++k) {
        const Plan& q = plan[k];
        if (q.mode != DWBA_PREP_FILON || q.N1 < 2) continue;
        auto key = std::make_tuple(q.energy, q.N1, q.log1, q.log2, q.lay.p_start, q.lay.p_step, q.lay.p_end,
                                   q.lay.n_pnodes, (int)q.lay.coulomb, q.lay.coulomb ? q.lay.k_total : 0.0,
                                   q.lay.coulomb ? q.lay.eta_total : 0.0);
        auto it = gkey.find(key);
        if (it == gkey.end()) {
            it = gkey.emplace(key, (int)groups.size()).first;
            groups.emplace_back();
        }
        groups[it->second].push_back(k);
    }
    // the fused pass's slices (<= GROUP_SLICE members of a group): each
    // energy's slices in one XCD run of dwba_filon_fused_f64 (its grid
    // arrays, u_f and cell tables stay in one L2), energies spread over the
    // 8 runs by longest-processing-time, heaviest first in each, runs padded
    // to one length with empty slices (g < 0)
    struct SliceRef { int g, q, c; };
    std::vector<SliceRef> slice_list;
    {
        std::vector<int> e_order;
        std::map<int, std::vector<SliceRef>> per_e;
        std::map<int, double> e_cost;
        for (size_t g = 0; g < groups.size(); ++g) {
            const int e = plan[groups[g][0]].energy;
            if (!per_e.count(e)) e_order.push_back(e);
            const int cnt = (int)groups[g].size();
            for (int q = 0; q < cnt; q += GROUP_SLICE) {
                const int c = std::min(GROUP_SLICE, cnt - q);
                per_e[e].push_back({(int)g, q, c});
                e_cost[e] += (double)plan[groups[g][0]].N * (10 + 6 * c);
            }
        }
        std::stable_sort(e_order.begin(), e_order.end(), [&](int a, int b) { return e_cost[a] > e_cost[b]; });
        std::vector<SliceRef> bins[8];
        double load[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int e : e_order) {
            int b = 0;
            for (int i = 1; i < 8; ++i)
                if (load[i] < load[b]) b = i;
            bins[b].insert(bins[b].end(), per_e[e].begin(), per_e[e].end());
            load[b] += e_cost[e];
        }
        size_t len = 0;
        for (auto& b : bins) len = std::max(len, b.size());
        for (auto& b : bins) {
            slice_list.insert(slice_list.end(), b.begin(), b.end());
            for (size_t i = b.size(); i < len; ++i) slice_list.push_back({-1, 0, 0});
        }
    }
    const long n_slices = (long)slice_list.size();
    const bool grouped = !groups.empty() && n_slices <= 65535;
    std::vector<char> in_group(n_pairs, 0);
    if (grouped)
        for (auto& g : groups)
            for (int k : g) in_group[k] = 1;

    // ---- moment sets (energy, u_f) and tail jobs, longest first
    std::vector<int> tail_ids;
    for (int k = 0; k < n_pairs; ++k)
        if (plan[k].tail != DWBA_TAIL_NONE) tail_ids.push_back(k);
    filon_limits(s->max_seg, s->max_dphi);
    {
        std::vector<long> len(n_pairs, 0);
        for (int k : tail_ids) {
            const Plan& q = plan[k];
            len[k] = tail_sweep_nodes(q.k_i, q.l_i, q.eta_i, q.k_f, q.l_f, q.eta_f, q.r_match, q.r_max, s->max_seg,
                                      s->max_dphi);
        }
        std::stable_sort(tail_ids.begin(), tail_ids.end(), [&](int a, int b) { return len[a] > len[b]; });
    }
    std::map<std::pair<int, const double*>, int> set_of;
    std::vector<std::pair<int, const double*>> sets;
    for (int k : tail_ids) {
        auto key = std::make_pair(P[k].energy, P[k].u_f);
        if (!set_of.count(key)) {
            set_of[key] = (int)sets.size();
            sets.push_back(key);
        }
    }
    levin_tables(s->levin_D, s->levin_x);

    // ---- arena layout
    Arena A;
    std::vector<size_t> o_r(n_energies), o_lr(n_energies), o_w(n_energies), o_vd(n_energies), o_ui(n_energies),
        o_tab(n_energies);
    for (int e = 0; e < n_energies; ++e) {
        const size_t nb = sizeof(double) * E[e].n;
        o_r[e] = A.take(nb);
        o_lr[e] = A.take(nb);
        o_w[e] = A.take(nb);
        o_vd[e] = A.take(nb);
        o_ui[e] = A.take(nb);
        o_tab[e] = A.take(nb * stride);
    }
    std::vector<size_t> o_chi(n_waves, 0);
    for (int w = 0; w < n_waves; ++w)
        if (wave_energy[w] >= 0) o_chi[w] = A.take(sizeof(double) * E[wave_energy[w]].n);
    std::map<std::pair<int, const double*>, size_t> o_uf;
    for (int k = 0; k < n_pairs; ++k) {
        auto key = std::make_pair(P[k].energy, P[k].u_f);
        if (!o_uf.count(key)) o_uf[key] = A.take(sizeof(double) * E[P[k].energy].n);
    }
    std::vector<size_t> o_vec(n_pairs);
    for (int k = 0; k < n_pairs; ++k) o_vec[k] = A.take(sizeof(double) * 4 * ld[P[k].energy]);
    const size_t o_scal = A.take(sizeof(double) * 4 * n_pairs);
    std::vector<size_t> o_tails(n_pairs, 0);
    for (int k : tail_ids) o_tails[k] = A.take(sizeof(double) * nl);
    // Coulomb phase nodes: one buffer per distinct stepping of an energy
    std::map<std::tuple<int, double, double, double, double, double>, std::pair<size_t, int>> node_owner;
    std::vector<size_t> o_nodes(n_pairs, 0);
    std::vector<char> nodes_shared(n_pairs, 0);
    for (int k = 0; k < n_pairs; ++k) {
        const Plan& q = plan[k];
        if (q.mode != DWBA_PREP_FILON || !q.lay.coulomb) continue;
        auto key = std::make_tuple(q.energy, q.lay.p_start, q.lay.p_end, q.lay.p_step, q.lay.k_total, q.lay.eta_total);
        auto it = node_owner.find(key);
        if (it == node_owner.end()) {
            it = node_owner.emplace(key, std::make_pair(A.take(sizeof(double) * COUL_NODES), k)).first;
        } else {
            nodes_shared[k] = 1;
        }
        o_nodes[k] = it->second.first;
    }
    const bool any_coul = !node_owner.empty();
    const size_t o_coul = any_coul ? A.take(sizeof(double) * 4 * n_pairs) : 0;
    const size_t o_prep = A.take(sizeof(dwba_prep_job) * n_pairs);
    size_t o_coef_total = 0;
    std::vector<int> ld_coef(groups.size());
    std::vector<size_t> o_coef(groups.size());
    if (grouped) {
        for (size_t g = 0; g < groups.size(); ++g) {
            ld_coef[g] = (plan[groups[g][0]].N1 + 31) / 32 * 32;
            o_coef[g] = A.take(sizeof(double) * 7 * ld_coef[g]);
            o_coef_total += 7 * ld_coef[g];
        }
    }
    const size_t o_groups = grouped ? A.take(sizeof(dwba_filon_group) * groups.size()) : 0;
    const size_t o_slices = grouped ? A.take(sizeof(dwba_filon_group) * n_slices) : 0;
    const size_t o_members = grouped ? A.take(sizeof(int32_t) * n_pairs) : 0;
    const int ns = (int)sets.size();
    const size_t o_mom = ns ? A.take(sizeof(double) * ns * nl) : 0;
    const size_t o_momptr = ns ? A.take(sizeof(int64_t) * 4 * ns) : 0;
    const size_t o_momn = ns ? A.take(sizeof(int32_t) * ns) : 0;
    const int nt = (int)tail_ids.size();
    const size_t tail_ws_bytes = nt ? dwba_outer_tail_workspace_bytes(nt, Lm) : 0;
    const size_t o_tj = nt ? A.take(sizeof(dwba_tail_job) * nt) : 0;
    const size_t o_tws = nt ? A.take(tail_ws_bytes) : 0;
    const size_t o_lgrid = A.take(sizeof(dwba_lane_grid) * n_energies);
    const size_t o_jobs = A.take(sizeof(dwba_plain_job) * n_slots);
    const size_t o_tabs = A.take(sizeof(int64_t) * n_slots);
    const size_t o_oidx = A.take(sizeof(int32_t) * n_slots);
    const int ng = (int)gemv_energies.size();
    std::vector<size_t> o_t(ng);
    for (int g = 0; g < ng; ++g) o_t[g] = A.take(sizeof(double) * nl * ld[gemv_energies[g]]);
    int n_dots = 0, n_items = 0;
    for (auto& ks : gemv_pairs) {
        n_dots += (int)ks.size();
        n_items += ((int)ks.size() + 15) / 16;
    }
    const size_t o_gemv = ng ? A.take(sizeof(dwba_dgemv_energy) * ng) : 0;
    const size_t o_dots = ng ? A.take(sizeof(dwba_ddot_pair) * n_dots) : 0;
    const size_t o_items = ng ? A.take(sizeof(int32_t) * 2 * n_items) : 0;
    const size_t o_fin = A.take(sizeof(dwba_finalize_pair) * n_pairs);
    const size_t o_out = A.take(sizeof(double) * 2 * nl * n_pairs);

    // ---- allocation (zeroed once: the Filon vectors keep their zeros beyond
    // the points the quadrature nodes touch on every launch)
    s->bytes = A.top;
    DWBA_HIP_CHECK(hipMalloc(&s->base, s->bytes));
    DWBA_HIP_CHECK(hipMemset(s->base, 0, s->bytes));
    char* B = static_cast<char*>(s->base);
    auto dptr = [&](size_t off) { return reinterpret_cast<double*>(B + off); };
    auto up = [&](size_t off, const void* src, size_t nb) { return hipMemcpy(B + off, src, nb, hipMemcpyHostToDevice); };

    // ---- uploads
    for (int e = 0; e < n_energies; ++e) {
        const dwba_scan_energy& en = E[e];
        const int N = en.n;
        std::vector<double> tmp(N);
        DWBA_HIP_CHECK(up(o_r[e], en.r, sizeof(double) * N));
        for (int i = 0; i < N; ++i) tmp[i] = std::log(std::max(en.r[i], 1e-300));
        DWBA_HIP_CHECK(up(o_lr[e], tmp.data(), sizeof(double) * N));
        DWBA_HIP_CHECK(up(o_w[e], en.w_simpson, sizeof(double) * N));
        for (int i = 0; i < N; ++i) tmp[i] = en.V_core[i] - en.U_i[i];
        DWBA_HIP_CHECK(up(o_vd[e], tmp.data(), sizeof(double) * N));
        DWBA_HIP_CHECK(up(o_ui[e], en.u_i, sizeof(double) * N));
    }
    for (int w = 0; w < n_waves; ++w)
        if (wave_energy[w] >= 0) DWBA_HIP_CHECK(up(o_chi[w], W[w].chi, sizeof(double) * E[wave_energy[w]].n));
    for (auto& kv : o_uf) DWBA_HIP_CHECK(up(kv.second, kv.first.second, sizeof(double) * E[kv.first.first].n));

    // ---- prep jobs (FILON pairs whose densities are finite beyond N1 let the
    // fused pass stop there)
    std::vector<int> nf_grid(n_energies), nf_wave(n_waves, -1);
    for (int e = 0; e < n_energies; ++e) {
        int v = last_nonfinite(E[e].u_i, E[e].n);
        for (int i = E[e].n - 1; i > v; --i)
            if (!std::isfinite(E[e].V_core[i] - E[e].U_i[i])) { v = i; break; }
        nf_grid[e] = v;
    }
    for (int w = 0; w < n_waves; ++w)
        if (wave_energy[w] >= 0) nf_wave[w] = last_nonfinite(W[w].chi, E[wave_energy[w]].n);
    std::map<std::pair<int, const double*>, int> nf_uf;
    for (auto& kv : o_uf) nf_uf[kv.first] = last_nonfinite(kv.first.second, E[kv.first.first].n);
    std::vector<dwba_prep_job> pj(n_pairs);
    std::vector<double> coul(any_coul ? 4 * n_pairs : 0, 0.0);
    for (int k = 0; k < n_pairs; ++k) {
        const Plan& q = plan[k];
        const int e = q.energy, L = ld[e];
        dwba_prep_job& J = pj[k];
        std::memset(&J, 0, sizeof(J));
        J.r = dptr(o_r[e]);
        J.logr = dptr(o_lr[e]);
        J.w = dptr(o_w[e]);
        J.u_i = dptr(o_ui[e]);
        J.u_f = dptr(o_uf[{e, P[k].u_f}]);
        J.chi_i = dptr(o_chi[P[k].wave_i]);
        J.chi_f = dptr(o_chi[P[k].wave_f]);
        J.V_diff = dptr(o_vd[e]);
        J.phase_nodes = o_nodes[k] ? dptr(o_nodes[k]) : nullptr;
        double* v = dptr(o_vec[k]);
        J.alpha_dir = v;
        J.beta_dir = no_beta[k] ? nullptr : v + L;
        J.alpha_ex = v + 2 * L;
        J.beta_ex = v + 3 * L;
        J.scalars = dptr(o_scal) + 4 * k;
        J.p_start = q.lay.p_start;
        J.p_step = q.lay.p_step;
        J.p_end = q.lay.p_end;
        J.N = q.N;
        J.N1 = q.N1;
        J.n_pnodes = q.lay.n_pnodes;
        int mode = q.mode;
        if (q.mode == DWBA_PREP_FILON) mode |= DWBA_PREP_NO_ZERO_FILL;
        if (nodes_shared[k]) mode |= DWBA_PREP_NODES_SHARED;
        if (q.mode == DWBA_PREP_FILON &&
            std::max(std::max(nf_grid[e], nf_uf[{e, P[k].u_f}]), std::max(nf_wave[P[k].wave_i], nf_wave[P[k].wave_f])) <
                q.N1)
            mode |= DWBA_PREP_FINITE_TAIL;
        if (in_group[k]) mode |= DWBA_PREP_FILON_GROUPED;
        J.mode = mode;
        J.log1 = q.log1;
        J.log2 = q.log2;
        if (q.mode == DWBA_PREP_FILON && q.lay.coulomb) {
            coul[4 * k + 0] = q.lay.p_start;
            coul[4 * k + 1] = q.lay.p_end;
            coul[4 * k + 2] = q.lay.k_total;
            coul[4 * k + 3] = q.lay.eta_total;
        }
    }
    DWBA_HIP_CHECK(up(o_prep, pj.data(), sizeof(dwba_prep_job) * n_pairs));
    s->prep = at<dwba_prep_job>(s->base, o_prep);
    if (any_coul) {
        DWBA_HIP_CHECK(up(o_coul, coul.data(), sizeof(double) * coul.size()));
        s->coul = dptr(o_coul);
        s->any_coul = 1;
    }

    // ---- Filon groups and the fused pass's slices
    if (grouped) {
        std::vector<dwba_filon_group> ga(groups.size());
        std::vector<int32_t> mem;
        int first = 0;
        for (size_t g = 0; g < groups.size(); ++g) {
            ga[g].coef = dptr(o_coef[g]);
            ga[g].ld_coef = ld_coef[g];
            ga[g].rep = groups[g][0];
            ga[g].first = first;
            ga[g].count = (int)groups[g].size();
            for (int k : groups[g]) mem.push_back(k);
            first += (int)groups[g].size();
        }
        std::vector<dwba_filon_group> sl;
        for (const SliceRef& r : slice_list) {
            dwba_filon_group x = ga[r.g >= 0 ? r.g : 0];
            x.first = r.g >= 0 ? x.first + r.q : 0;
            x.count = r.c;
            sl.push_back(x);
        }
        DWBA_HIP_CHECK(up(o_groups, ga.data(), sizeof(dwba_filon_group) * ga.size()));
        DWBA_HIP_CHECK(up(o_slices, sl.data(), sizeof(dwba_filon_group) * sl.size()));
        DWBA_HIP_CHECK(up(o_members, mem.data(), sizeof(int32_t) * mem.size()));
        s->groups = at<dwba_filon_group>(s->base, o_groups);
        s->slices = at<dwba_filon_group>(s->base, o_slices);
        s->members = at<int32_t>(s->base, o_members);
        s->n_groups = (int32_t)ga.size();
        s->n_slices = (int32_t)sl.size();
    }

    // ---- moments and tails
    if (nt) {
        std::vector<int64_t> mp(4 * ns);
        std::vector<int32_t> mn(ns);
        for (int q = 0; q < ns; ++q) {
            const int e = sets[q].first;
            mp[0 * ns + q] = (int64_t)(uintptr_t)dptr(o_r[e]);
            mp[1 * ns + q] = (int64_t)(uintptr_t)dptr(o_w[e]);
            mp[2 * ns + q] = (int64_t)(uintptr_t)dptr(o_ui[e]);
            mp[3 * ns + q] = (int64_t)(uintptr_t)dptr(o_uf[sets[q]]);
            mn[q] = E[e].n;
        }
        DWBA_HIP_CHECK(up(o_momptr, mp.data(), sizeof(int64_t) * mp.size()));
        DWBA_HIP_CHECK(up(o_momn, mn.data(), sizeof(int32_t) * ns));
        for (int c = 0; c < 4; ++c) s->mom_ptr[c] = at<const double* const>(s->base, o_momptr + sizeof(int64_t) * c * ns);
        s->mom_n = at<int32_t>(s->base, o_momn);
        s->moments = dptr(o_mom);
        s->n_sets = ns;
        std::vector<dwba_tail_job> tj(nt);
        for (int q = 0; q < nt; ++q) {
            const int k = tail_ids[q];
            const Plan& p = plan[k];
            dwba_tail_job& T = tj[q];
            std::memset(&T, 0, sizeof(T));
            T.k_a = p.k_i;
            T.delta_a = p.delta_i;
            T.eta_a = p.eta_i;
            T.sigma_a = p.sigma_i;
            T.k_b = p.k_f;
            T.delta_b = p.delta_f;
            T.eta_b = p.eta_f;
            T.sigma_b = p.sigma_f;
            T.l_a = p.l_i;
            T.l_b = p.l_f;
            T.r_m = p.r_match;
            T.r_max = p.r_max;
            T.moments = dptr(o_mom) + (size_t)nl * set_of[{P[k].energy, P[k].u_f}];
            T.out = dptr(o_tails[k]);
            T.L_max = Lm;
            T.kind = p.tail;
        }
        DWBA_HIP_CHECK(up(o_tj, tj.data(), sizeof(dwba_tail_job) * nt));
        s->tail_jobs = at<dwba_tail_job>(s->base, o_tj);
        s->tail_ws = B + o_tws;
        s->tail_ws_bytes = tail_ws_bytes;
        s->n_tail = nt;
    }

    // ---- contraction: lane jobs (exchange, and direct where no shared beta),
    // pairs of one energy together, energies with the longest pairs first
    std::vector<dwba_plain_job> jobs(n_pairs);
    for (int k = 0; k < n_pairs; ++k) {
        const Plan& q = plan[k];
        const int e = q.energy, L = ld[e];
        double* v = dptr(o_vec[k]);
        dwba_plain_job& J = jobs[k];
        J.r = dptr(o_r[e]);
        const bool skip = shared_dir[k] != 0;
        J.alpha_dir = skip ? nullptr : v;
        J.beta_dir = skip ? nullptr : v + L;
        J.alpha_ex = v + 2 * L;
        J.beta_ex = v + 3 * L;
        J.n_rows = q.rows;
        // the exchange family alone stops at N1 for Filon / plain pairs (alpha_ex and beta_ex vanish beyond)
        J.n_cols = (skip && (q.mode == DWBA_PREP_FILON || q.mode == DWBA_PREP_PLAIN)) ? q.rows : q.cols;
    }
    std::vector<int> emax(n_energies, 0);
    for (int k = 0; k < n_pairs; ++k) emax[plan[k].energy] = std::max(emax[plan[k].energy], jobs[k].n_cols);
    std::vector<int> perm(n_pairs);
    std::iota(perm.begin(), perm.end(), 0);
    std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) {
        const int ea = plan[a].energy, eb = plan[b].energy;
        if (emax[ea] != emax[eb]) return emax[ea] > emax[eb];
        if (ea != eb) return ea < eb;
        return jobs[a].n_cols > jobs[b].n_cols;
    });
    {
        // energy order (energies with the longest jobs first, longest first within each); for the staged
        // kernel each energy's jobs are padded to whole groups by empty jobs (output index -1)
        std::vector<dwba_plain_job> sj;
        std::vector<int64_t> tp;
        std::vector<int32_t> oi;
        sj.reserve(n_slots);
        for (int p = 0; p < n_pairs; ++p) {
            const int e = plan[perm[p]].energy;
            sj.push_back(jobs[perm[p]]);
            tp.push_back((int64_t)(uintptr_t)dptr(o_tab[e]));
            oi.push_back(perm[p]);
            const bool last = p + 1 == n_pairs || plan[perm[p + 1]].energy != e;
            while (s->stage && last && sj.size() % stage_group) {
                dwba_plain_job z = jobs[perm[p]];
                z.n_rows = 0;
                z.n_cols = 0;
                sj.push_back(z);
                tp.push_back(tp.back());
                oi.push_back(-1);
            }
        }
        DWBA_REQUIRE((int)sj.size() <= n_slots, "dwba_scan_plan: %zu job slots > %d", sj.size(), n_slots);
        DWBA_HIP_CHECK(up(o_jobs, sj.data(), sizeof(dwba_plain_job) * sj.size()));
        DWBA_HIP_CHECK(up(o_tabs, tp.data(), sizeof(int64_t) * tp.size()));
        DWBA_HIP_CHECK(up(o_oidx, oi.data(), sizeof(int32_t) * oi.size()));
        s->jobs = at<dwba_plain_job>(s->base, o_jobs);
        s->tabs = at<const double* const>(s->base, o_tabs);
        s->out_index = at<int32_t>(s->base, o_oidx);
        s->n_jobs = (int32_t)sj.size();
    }
    if (ng) {
        std::vector<dwba_dgemv_energy> dg(ng);
        std::vector<dwba_ddot_pair> dd;
        std::vector<int32_t> items;
        for (int g = 0; g < ng; ++g) {
            const int e = gemv_energies[g];
            dg[g].r = dptr(o_r[e]);
            dg[g].beta = dptr(o_vec[gemv_owner[g]]) + ld[e];
            dg[g].t = dptr(o_t[g]);
            dg[g].N = E[e].n;
            dg[g].rows = gemv_rows[g];
            dg[g].ld = ld[e];
            dg[g].L_max = Lm;
            const int first = (int)dd.size();
            for (int k : gemv_pairs[g]) {
                dwba_ddot_pair d;
                std::memset(&d, 0, sizeof(d));
                d.alpha = dptr(o_vec[k]);
                d.t = dptr(o_t[g]);
                d.n = plan[k].N1;
                d.ld = ld[e];
                d.out_index = k;
                dd.push_back(d);
            }
            const int cnt = (int)gemv_pairs[g].size();
            for (int q = 0; q < cnt; q += 16) {
                items.push_back(first + q);
                items.push_back(std::min(16, cnt - q));
            }
        }
        DWBA_HIP_CHECK(up(o_gemv, dg.data(), sizeof(dwba_dgemv_energy) * ng));
        DWBA_HIP_CHECK(up(o_dots, dd.data(), sizeof(dwba_ddot_pair) * dd.size()));
        DWBA_HIP_CHECK(up(o_items, items.data(), sizeof(int32_t) * items.size()));
        s->gemv = at<dwba_dgemv_energy>(s->base, o_gemv);
        s->dots = at<dwba_ddot_pair>(s->base, o_dots);
        s->items16 = at<int32_t>(s->base, o_items);
        s->n_gemv = ng;
        s->n_dots = (int32_t)dd.size();
        s->n_items16 = (int32_t)(items.size() / 2);
    }

    // ---- finalize table
    {
        std::vector<dwba_finalize_pair> fp(n_pairs);
        for (int k = 0; k < n_pairs; ++k) {
            std::memset(&fp[k], 0, sizeof(fp[k]));
            fp[k].scalars = dptr(o_scal) + 4 * k;
            fp[k].tails = o_tails[k] ? dptr(o_tails[k]) : nullptr;
            fp[k].out_index = k;
            fp[k].oscillatory = plan[k].osc ? 1 : 0;
            fp[k].tail_kind = plan[k].tail;
        }
        DWBA_HIP_CHECK(up(o_fin, fp.data(), sizeof(dwba_finalize_pair) * n_pairs));
        s->fin = at<dwba_finalize_pair>(s->base, o_fin);
    }
    s->out = dptr(o_out);

    // ---- streams, events; the per-grid lane tables (they depend only on r)
    DWBA_HIP_CHECK(hipDeviceSynchronize());
    DWBA_HIP_CHECK(hipStreamCreateWithFlags(&s->side, hipStreamNonBlocking));
    DWBA_HIP_CHECK(hipEventCreateWithFlags(&s->ev_fork, hipEventDisableTiming));
    DWBA_HIP_CHECK(hipEventCreateWithFlags(&s->ev_join, hipEventDisableTiming));
    {
        std::vector<dwba_lane_grid> lg(n_energies);
        for (int e = 0; e < n_energies; ++e) {
            lg[e].r = dptr(o_r[e]);
            lg[e].tab = dptr(o_tab[e]);
            lg[e].n = E[e].n;
            lg[e].pad_ = 0;
        }
        DWBA_HIP_CHECK(up(o_lgrid, lg.data(), sizeof(dwba_lane_grid) * n_energies));
        const int rc = dwba_semisep_lane_tables_f64(n_energies, at<dwba_lane_grid>(s->base, o_lgrid), max_n, Lm,
                                                    s->lane_lb, s->side);
        if (rc) return rc;
        DWBA_HIP_CHECK(hipStreamSynchronize(s->side));
    }
    s->info[0] = n_pairs;
    s->info[1] = Lm;
    s->info[2] = (int64_t)s->bytes;
    s->info[3] = nt;
    s->info[4] = s->n_groups;
    s->info[5] = ng;
    s->info[6] = (int64_t)node_owner.size();
    s->info[7] = ns;
    (void)o_coef_total;
    return DWBA_OK;
}

}  // namespace

extern "C" {

int32_t dwba_scan_plan(int32_t n_energies, const dwba_scan_energy* h_energies, int32_t n_waves,
                       const dwba_scan_wave* h_waves, int32_t n_pairs, const dwba_scan_pair* h_pairs,
                       const dwba_scan_options* h_options, dwba_scan** out_scan) {
    DWBA_REQUIRE(dwba::initialised(), "dwba_scan_plan: call dwba_init first");
    DWBA_REQUIRE(out_scan, "dwba_scan_plan: out_scan is NULL");
    *out_scan = nullptr;
    DWBA_REQUIRE(n_energies > 0 && h_energies && n_waves > 0 && h_waves && n_pairs > 0 && h_pairs && h_options,
                 "dwba_scan_plan: bad arguments (n_energies=%d, n_waves=%d, n_pairs=%d)", n_energies, n_waves, n_pairs);
    dwba_scan* s = new (std::nothrow) dwba_scan();
    DWBA_REQUIRE(s, "dwba_scan_plan: out of host memory");
    (void)hipGetDevice(&s->device);
    const int rc = plan_scan(s, n_energies, h_energies, n_waves, h_waves, n_pairs, h_pairs, h_options);
    if (rc != DWBA_OK) {
        const std::string msg = dwba::last_error();
        scan_free(s);
        dwba::last_error() = msg;
        return rc;
    }
    *out_scan = s;
    return DWBA_OK;
}

int32_t dwba_scan_plan_pairs(int32_t n_energies, const dwba_scan_energy* h_energies, int32_t n_waves,
                             const dwba_scan_wave* h_waves, int32_t n_pairs, const dwba_scan_pair* h_pairs,
                             const dwba_scan_options* h_options, dwba_scan_pair_plan* h_out) {
    DWBA_REQUIRE(n_energies > 0 && h_energies && n_waves > 0 && h_waves && n_pairs > 0 && h_pairs && h_options && h_out,
                 "dwba_scan_plan_pairs: bad arguments (n_energies=%d, n_waves=%d, n_pairs=%d)", n_energies, n_waves,
                 n_pairs);
    std::vector<GridStats> gs;
    std::vector<int> wave_energy;
    std::vector<Plan> plan;
    const int rc = make_plans("dwba_scan_plan_pairs", n_energies, h_energies, n_waves, h_waves, n_pairs, h_pairs,
                              h_options, gs, wave_energy, plan);
    if (rc) return rc;
    for (int k = 0; k < n_pairs; ++k) {
        const Plan& q = plan[k];
        dwba_scan_pair_plan& o = h_out[k];
        std::memset(&o, 0, sizeof(o));
        o.p_start = q.lay.p_start;
        o.p_step = q.lay.p_step;
        o.p_end = q.lay.p_end;
        o.r_match = q.r_match;
        o.N1 = q.N1;
        o.mode = q.mode;
        o.tail = q.tail;
        o.rows = q.rows;
        o.cols = q.cols;
        o.log1 = q.log1;
        o.log2 = q.log2;
        o.n_pnodes = q.lay.n_pnodes;
        o.coulomb = q.lay.coulomb ? 1 : 0;
        o.oscillatory = q.osc ? 1 : 0;
    }
    return DWBA_OK;
}

int32_t dwba_scan_launch(dwba_scan* scan, void* stream) {
    DWBA_REQUIRE(scan, "dwba_scan_launch: scan is NULL");
    return enqueue(scan, reinterpret_cast<hipStream_t>(stream));
}

int32_t dwba_scan_capture(dwba_scan* scan, void* stream) {
    DWBA_REQUIRE(scan, "dwba_scan_capture: scan is NULL");
    DWBA_REQUIRE(stream, "dwba_scan_capture: capture needs a created stream, not the null stream");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if (scan->exec) {
        DWBA_HIP_CHECK(hipGraphExecDestroy(scan->exec));
        scan->exec = nullptr;
    }
    if (scan->graph) {
        DWBA_HIP_CHECK(hipGraphDestroy(scan->graph));
        scan->graph = nullptr;
    }
    DWBA_HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue(scan, st);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(st, &g);
    if (rc != DWBA_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return dwba::set_error(DWBA_ERR_HIP, "dwba_scan_capture: hipStreamEndCapture failed: %s", hipGetErrorString(e));
    }
    scan->graph = g;
    DWBA_HIP_CHECK(hipGraphInstantiate(&scan->exec, g, nullptr, nullptr, 0));
    return DWBA_OK;
}

int32_t dwba_scan_replay(dwba_scan* scan, void* stream) {
    DWBA_REQUIRE(scan && scan->exec, "dwba_scan_replay: no captured graph (call dwba_scan_capture first)");
    DWBA_HIP_CHECK(hipGraphLaunch(scan->exec, reinterpret_cast<hipStream_t>(stream)));
    return DWBA_OK;
}

const double* dwba_scan_results(const dwba_scan* scan) { return scan ? scan->out : nullptr; }

int32_t dwba_scan_read_results(const dwba_scan* scan, double* h_out, void* stream) {
    DWBA_REQUIRE(scan && h_out, "dwba_scan_read_results: bad arguments");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const size_t nb = sizeof(double) * 2 * (size_t)(scan->L_max + 1) * scan->n_pairs;
    DWBA_HIP_CHECK(hipMemcpyAsync(h_out, scan->out, nb, hipMemcpyDeviceToHost, st));
    DWBA_HIP_CHECK(hipStreamSynchronize(st));
    return DWBA_OK;
}

int32_t dwba_scan_info(const dwba_scan* scan, int64_t* h_info, int32_t n_info) {
    DWBA_REQUIRE(scan && h_info && n_info >= 0, "dwba_scan_info: bad arguments");
    for (int i = 0; i < n_info; ++i) h_info[i] = i < 8 ? scan->info[i] : 0;
    return DWBA_OK;
}

int32_t dwba_scan_destroy(dwba_scan* scan) {
    if (scan) {
        (void)hipDeviceSynchronize();
        scan_free(scan);
    }
    return DWBA_OK;
}

}  // extern "C"
