"""Oracle: radial_ME_all_L (direct + exchange I_L for L = 0..L_max).

Test infrastructure only.  Two independent routes to the same numbers:

* ``radial_ME_all_L_faithful`` restates dwba_matrix_elements.py:780-1300 and
  oscillatory_kernel_integral_2d (oscillatory_integrals.py:2385-2780) step by
  step, materialising the N x N kernel K_L for every L exactly as the
  reference does (O(N^2) memory, O(L N^2) work).  It is also the CPU baseline
  that bench.py times.

* ``radial_ME_all_L_reduced`` rewrites every quadrature branch as
  I_L = alpha^T K_L beta (the Clenshaw-Curtis node interpolations are linear
  in the densities), evaluates the contraction with the exact O(N) long
  double form (semisep.py), and adds the same tails and L=0 corrections.

Agreement of the two routes pins the alpha/beta reduction that the GPU path
implements.  Input objects are duck-typed like the reference's:
``bound.u_of_r``; ``cont.chi_of_r, k_au, l, idx_match, phase_shift, eta,
sigma_l``.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Dict, Optional

import numpy as np

from . import quadrature as Q
from . import semisep
from . import tail as T


@dataclass
class Result:
    I_L_direct: Dict[int, float]
    I_L_exchange: Dict[int, float]


def _attr(obj, name, default):
    return getattr(obj, name) if hasattr(obj, name) else default


# --------------------------------------------------------------------------
# split index (dwba_matrix_elements.py:909-1022)
# --------------------------------------------------------------------------
def compute_idx_limit(r, w_trapz, U_i, U_f, u_i, u_f, k_i, k_f, l_i, l_f,
                      idx_match_i, idx_match_f, min_grid_fraction=0.10) -> int:
    n = len(r)
    idx_limit = n
    if 0 < idx_match_i < n:
        idx_limit = min(idx_limit, idx_match_i + 1)
    if 0 < idx_match_f < n:
        idx_limit = min(idx_limit, idx_match_f + 1)
    k_eff = min(k_i, k_f) if min(k_i, k_f) > 0.1 else max(k_i, k_f)
    l_max_wave = max(l_i, l_f)
    if k_eff > 1e-3:
        r_turn = (l_max_wave + 0.5) / k_eff
        idx_turn = int(np.searchsorted(r, r_turn))
        idx_turn = min(idx_turn, n - 20)
        min_idx = min(max(idx_turn + 20, int(n * min_grid_fraction)), n)
    else:
        min_idx = max(1, int(n * min_grid_fraction))
    if idx_limit < min_idx:
        idx_limit = min_idx
    idx_99_max = 0
    for u in (u_i, u_f):
        u_sq = u ** 2
        if np.sum(u_sq) > 1e-30:
            pc = np.cumsum(u_sq * w_trapz)
            pc /= (pc[-1] + 1e-300)
            idx_99_max = max(idx_99_max, int(np.searchsorted(pc, 0.99)))
    if idx_99_max > idx_limit:
        idx_limit = idx_99_max + 1

    def veff(idx):
        ri = r[idx]
        ui = abs(U_i[idx])
        uf = abs(U_f[idx]) if U_f is not None else 0.0
        return max(ui, uf) + (l_max_wave * (l_max_wave + 1)) / (2.0 * ri ** 2)

    kin = 0.5 * min(k_i, k_f) ** 2
    if kin > 1e-10:
        if veff(max(0, idx_limit - 1)) / kin > 0.03:
            for t in range(idx_limit, min(idx_limit + 500, n)):
                if veff(t) / kin < 0.03:
                    idx_limit = t + 1
                    break
    return int(idx_limit)


# --------------------------------------------------------------------------
# faithful oscillatory_kernel_integral_2d (oscillatory_integrals.py:2385-2780)
# --------------------------------------------------------------------------
def _filon_nodes(r1_lim, k_total, phase_increment, eta_total, n_nodes):
    """Phase intervals and CC nodes shared by the filon and filon_exchange
    branches (oscillatory_integrals.py:2578-2628, :2667-2712).  Returns None
    when the branch falls back to the standard weighted dot product."""
    dphi = Q.adaptive_phase_increment(r1_lim, k_total, phase_increment, eta_total,
                                      threshold=np.pi / 4, min_increment=np.pi / 16)
    if k_total < 1e-6:
        return None
    pn = Q.generate_phase_nodes(r1_lim[0], r1_lim[-1], k_total, dphi, eta_total)
    if len(pn) - 1 < 1:
        return None
    a, b = pn[:-1], pn[1:]
    valid = (b - a) > 1e-12
    a, b = a[valid], b[valid]
    if len(a) == 0:
        return None
    x_ref, w_ref = Q.cc_ref(n_nodes)
    hw = 0.5 * (b - a)
    all_r = hw[:, None] * (x_ref + 1) + a[:, None]
    W = w_ref * hw[:, None]
    return all_r, W


def kernel_integral_2d(rho1, rho2, kernel, r, k_i, k_f, idx_limit=-1, idx_limit_r2=None,
                       method="filon", n_nodes=5, phase_increment=np.pi / 2, eta_total=0.0,
                       w_grid=None) -> float:
    if idx_limit < 0:
        idx_limit = len(r)
    if idx_limit_r2 is None or idx_limit_r2 < 0:
        idx_limit_r2 = len(r)
    if idx_limit < 2 or idx_limit_r2 < 2:
        return 0.0
    rho1_lim = rho1[:idx_limit].copy()
    rho2_lim = rho2[:idx_limit_r2].copy()
    K = kernel[:idx_limit, :idx_limit_r2]
    r1 = r[:idx_limit]
    r2 = r[:idx_limit_r2]
    w1 = w_grid[:idx_limit].copy() if w_grid is not None else np.ones(idx_limit)
    w2 = w_grid[:idx_limit_r2].copy() if w_grid is not None else np.ones(idx_limit_r2)
    if not np.all(np.isfinite(rho1_lim)) or not np.all(np.isfinite(rho2_lim)):
        rho1_lim = np.nan_to_num(rho1_lim, nan=0.0, posinf=0.0, neginf=0.0)
        rho2_lim = np.nan_to_num(rho2_lim, nan=0.0, posinf=0.0, neginf=0.0)
    k_total = k_i + k_f
    if method == "standard":
        res = np.dot(rho1_lim * w1, np.dot(K, rho2_lim * w2))
        return float(res) if np.isfinite(res) else 0.0
    if method == "filon":
        int_r2 = np.dot(K, rho2_lim * w2)
        if not np.all(np.isfinite(int_r2)):
            int_r2 = np.nan_to_num(int_r2, nan=0.0, posinf=0.0, neginf=0.0)
        nodes = _filon_nodes(r1, k_total, phase_increment, eta_total, n_nodes)
        if nodes is None:
            res = float(np.dot(rho1_lim * w1, int_r2))
        else:
            all_r, W = nodes
            flat = all_r.ravel()
            integrand = (Q.interp_on_radial_grid(flat, r1, rho1_lim)
                         * Q.interp_on_radial_grid(flat, r1, int_r2)).reshape(W.shape)
            res = float(np.sum(integrand * W))
        return float(res) if np.isfinite(res) else 0.0
    if method == "filon_exchange":
        nodes = _filon_nodes(r1, k_total, phase_increment, eta_total, n_nodes)
        if nodes is None:
            res = float(np.dot(rho1_lim * w1, np.dot(K, rho2_lim * w2)))
        else:
            all_r, W = nodes
            flat = all_r.ravel()
            n_valid, nn = W.shape
            rho2_cc = Q.interp_on_radial_grid(flat, r2, rho2_lim).reshape(n_valid, nn)
            ir = np.searchsorted(r2, flat)
            ir = np.clip(ir, 1, len(r2) - 1)
            il = ir - 1
            wr = (flat - r2[il]) / (r2[ir] - r2[il] + 1e-30)
            wl = 1.0 - wr
            k_cc = (K[:, il] * wl + K[:, ir] * wr).reshape(len(r1), n_valid, nn)
            int_r2_cc = np.sum(k_cc * rho2_cc[None] * W[None], axis=(1, 2))
            outer = (Q.interp_on_radial_grid(flat, r1, rho1_lim)
                     * Q.interp_on_radial_grid(flat, r1, int_r2_cc)).reshape(n_valid, nn)
            res = float(np.sum(outer * W))
        return float(res) if np.isfinite(res) else 0.0
    raise ValueError(f"Unknown method: {method}")


# --------------------------------------------------------------------------
# shared pieces of radial_ME_all_L
# --------------------------------------------------------------------------
class _Setup:
    """Everything radial_ME_all_L computes before its L loop
    (dwba_matrix_elements.py:884-1108)."""

    def __init__(self, r, w_simpson, w_trapz, V_core, U_i, bound_i, bound_f, cont_i, cont_f,
                 L_max, use_osc, method, CC_nodes, phase_increment, min_grid_fraction,
                 k_threshold, U_f, is_excitation):
        if L_max < 0:
            raise ValueError("radial_ME_all_L: L_max must be >= 0.")
        self.r = r = np.asarray(r, dtype=np.float64)
        self.N = len(r)
        self.w = w = np.asarray(w_simpson, dtype=np.float64)
        w_trapz = np.asarray(w_trapz, dtype=np.float64) if w_trapz is not None else np.gradient(r)
        self.u_i = u_i = np.asarray(bound_i.u_of_r, dtype=np.float64)
        self.u_f = u_f = np.asarray(bound_f.u_of_r, dtype=np.float64)
        self.k_i, self.k_f = k_i, k_f = float(cont_i.k_au), float(cont_f.k_au)
        self.l_i, self.l_f = int(cont_i.l), int(cont_f.l)
        chi_i = np.asarray(cont_i.chi_of_r, dtype=np.float64)
        chi_f = np.asarray(cont_f.chi_of_r, dtype=np.float64)
        self.idx_limit = compute_idx_limit(
            r, w_trapz, U_i, U_f, u_i, u_f, k_i, k_f, self.l_i, self.l_f,
            int(_attr(cont_i, "idx_match", -1)), int(_attr(cont_f, "idx_match", -1)),
            min_grid_fraction)
        wl = w.copy()
        wl[self.idx_limit:] = 0.0
        self.rho1_dir_w = wl * chi_f * chi_i
        self.rho2_dir_w = wl * u_f * u_i
        self.rho1_ex_w = wl * u_f * chi_i
        self.rho2_ex_w = wl * chi_f * u_i
        self.rho1_dir_uw = chi_f * chi_i
        self.rho2_dir_uw = u_f * u_i
        self.rho1_ex_uw = u_f * chi_i
        self.rho2_ex_uw = chi_f * u_i
        self.V_diff = np.asarray(V_core, dtype=np.float64) - np.asarray(U_i, dtype=np.float64)
        self.delta_i = float(_attr(cont_i, "phase_shift", 0.0))
        self.delta_f = float(_attr(cont_f, "phase_shift", 0.0))
        self.eta_i = float(_attr(cont_i, "eta", 0.0))
        self.eta_f = float(_attr(cont_f, "eta", 0.0))
        self.sigma_i = float(_attr(cont_i, "sigma_l", 0.0))
        self.sigma_f = float(_attr(cont_f, "sigma_l", 0.0))
        self.eta_total = self.eta_i + self.eta_f
        self.k_total = k_i + k_f
        self.r_match = r[self.idx_limit - 1] if self.idx_limit > 0 else r[-1]
        if is_excitation is None:
            is_excitation = not hasattr(bound_f, "k_au")
        self.is_excitation = bool(is_excitation)
        self.use_osc = bool(use_osc)
        self.method = method
        self.CC_nodes = int(CC_nodes)
        self.phase_increment = float(phase_increment)
        self.k_threshold = float(k_threshold)
        self.L_max = int(L_max)
        self.oscillatory = self.use_osc and self.k_total > self.k_threshold

    def moment(self, L):
        """moment_L = sum(w r^L u_f u_i) on the full grid (dwba_matrix_elements.py:1167)."""
        return np.sum(self.w * (self.r ** L) * self.u_f * self.u_i)

    def direct_tail(self, L) -> float:
        """I_out for full_split/advanced (dwba_matrix_elements.py:1165-1192,
        :1216-1241): outer Filon/Levin integral on [r_m, r_max] plus the
        analytical tail beyond r_max."""
        if not (L >= 1 and self.idx_limit < self.N - 10 and self.is_excitation):
            return 0.0
        m = self.moment(L)
        if abs(m) <= 1e-12:
            return 0.0
        env = T.multipole_envelope(m, L)
        out = T.dwba_outer_integral(env, self.k_i, self.l_i, self.delta_i, self.eta_i, self.sigma_i,
                                    self.k_f, self.l_f, self.delta_f, self.eta_f, self.sigma_f,
                                    self.r_match, self.r[-1], delta_phi=np.pi / 4)
        out += T.analytical_multipole_tail(self.r[-1], self.k_i, self.k_f, self.delta_i, self.delta_f,
                                           self.l_i, self.l_f, L, bound_overlap=m,
                                           eta_i=self.eta_i, eta_f=self.eta_f,
                                           sigma_i=self.sigma_i, sigma_f=self.sigma_f)
        return out

    def direct_tail_noise(self, L) -> float:
        """Conditioning noise of direct_tail (see tail.outer_integral_noise)."""
        if not (L >= 1 and self.idx_limit < self.N - 10 and self.is_excitation):
            return 0.0
        m = self.moment(L)
        if abs(m) <= 1e-12:
            return 0.0
        return T.outer_integral_noise(T.multipole_envelope(m, L), self.k_i, self.l_i, self.delta_i, self.eta_i,
                                      self.sigma_i, self.k_f, self.l_f, self.delta_f, self.eta_f, self.sigma_f,
                                      self.r_match, self.r[-1])

    def legacy_tail(self, L) -> float:
        """Legacy-method analytical tail at r_match (dwba_matrix_elements.py:1257-1268)."""
        if not (L >= 1 and self.use_osc and self.idx_limit < self.N - 10 and self.is_excitation):
            return 0.0
        m = self.moment(L)
        if abs(m) <= 1e-12:
            return 0.0
        return T.analytical_multipole_tail(self.r_match, self.k_i, self.k_f, self.delta_i,
                                           self.delta_f, self.l_i, self.l_f, L, m,
                                           self.eta_i, self.eta_f, self.sigma_i, self.sigma_f)

    def l0_correction(self, rho1_w, rho2_w) -> float:
        """L=0 term (dwba_matrix_elements.py:1249-1253, :1291-1296)."""
        s = np.sum(rho2_w)
        if abs(s) > 1e-12:
            return float(np.dot(rho1_w, self.V_diff) * s)
        return 0.0


def radial_ME_all_L_faithful(r, w_simpson, w_trapz, V_core, U_i, bound_i, bound_f, cont_i, cont_f,
                             L_max, use_oscillatory_quadrature=True, oscillatory_method="advanced",
                             CC_nodes=5, phase_increment=1.5708, min_grid_fraction=0.10,
                             k_threshold=0.5, U_f_array=None, is_excitation=None) -> Result:
    """Step-by-step restatement of radial_ME_all_L (dwba_matrix_elements.py:780-1300),
    materialising K_L (N x N) for every L."""
    s = _Setup(r, w_simpson, w_trapz, V_core, U_i, bound_i, bound_f, cont_i, cont_f, L_max,
               use_oscillatory_quadrature, oscillatory_method, CC_nodes, phase_increment,
               min_grid_fraction, k_threshold, U_f_array, is_excitation)
    r = s.r
    r1c, r2r = r[:, None], r[None, :]
    rl = np.minimum(r1c, r2r)
    rg = np.maximum(r1c, r2r)
    with np.errstate(divide="ignore", invalid="ignore"):
        ratio = rl / rg
        inv_gtr = 1.0 / rg
    ratio[~np.isfinite(ratio)] = 0.0
    inv_gtr[~np.isfinite(inv_gtr)] = 0.0
    log_ratio = np.log(np.minimum(ratio, 1.0 - 1e-12) + 1e-300)
    del rl, rg, ratio
    I_dir, I_ex = {}, {}
    for L in range(s.L_max + 1):
        K = inv_gtr if L == 0 else inv_gtr * np.exp(L * log_ratio)
        if s.oscillatory:
            i_in = kernel_integral_2d(s.rho1_dir_uw, s.rho2_dir_uw, K, r, s.k_i, s.k_f, s.idx_limit,
                                      None, "filon", s.CC_nodes, s.phase_increment, s.eta_total, s.w)
            if s.method == "legacy":
                d = i_in
            else:
                d = i_in + s.direct_tail(L)
        else:
            d = np.dot(s.rho1_dir_w, np.dot(K, s.rho2_dir_w))
        if L == 0:
            d += s.l0_correction(s.rho1_dir_w, s.rho2_dir_w)
        if s.method == "legacy":
            d += s.legacy_tail(L)
        I_dir[L] = float(d)
        if s.oscillatory:
            e = kernel_integral_2d(s.rho1_ex_uw, s.rho2_ex_uw, K, r, s.k_i, s.k_f, s.idx_limit,
                                   None, "filon_exchange", s.CC_nodes, s.phase_increment,
                                   s.eta_total, s.w)
        else:
            e = np.dot(s.rho1_ex_w, np.dot(K, s.rho2_ex_w))
        if L == 0:
            e += s.l0_correction(s.rho1_ex_w, s.rho2_ex_w)
        I_ex[L] = float(e)
    return Result(I_dir, I_ex)


# --------------------------------------------------------------------------
# alpha / beta reduction
# --------------------------------------------------------------------------
def _interp_weights(x_new, x):
    """Cell index j and weight t of _interp_on_radial_grid at each query, so
    that interp(y)(x_new) = (1-t) y[j] + t y[j+1] (np.interp semantics; at
    the last grid point j = N-2, t = 1)."""
    x = np.asarray(x, dtype=np.float64)
    xc = np.clip(np.asarray(x_new, dtype=np.float64), x[0], x[-1])
    if Q.is_log_spaced_grid(x):
        xp = np.log(np.maximum(x, 1e-300))
        q = np.log(np.maximum(xc, max(x[0], 1e-300)))
    else:
        xp, q = x, xc
    j = np.searchsorted(xp, q, side="right") - 1
    j = np.clip(j, 0, len(x) - 2)
    t = (q - xp[j]) / (xp[j + 1] - xp[j])
    t = np.clip(t, 0.0, 1.0)
    return j, t


def _scatter(j, t, v, n):
    out = np.zeros(n)
    np.add.at(out, j, v * (1.0 - t))
    np.add.at(out, j + 1, v * t)
    return out


def filon_alpha(rho1_uw, r, n1, nodes):
    """alpha of the CC outer quadrature: alpha_j = sum_c W_c interp(rho1)(x_c)
    * (interpolation weight of grid point j at node x_c), on r[:n1]."""
    all_r, W = nodes
    flat = all_r.ravel()
    r1 = r[:n1]
    j, t = _interp_weights(flat, r1)
    rho = np.nan_to_num(rho1_uw[:n1], nan=0.0, posinf=0.0, neginf=0.0)
    v = W.ravel() * ((1.0 - t) * rho[j] + t * rho[j + 1])
    return _scatter(j, t, v, len(r))


def filon_exchange_beta(rho2_uw, r, nodes):
    """beta of the exchange inner CC quadrature: node values of rho2 (log-r
    interpolation on the full grid) scattered onto the two grid columns that
    bracket the node with linear-in-r kernel interpolation weights."""
    all_r, W = nodes
    flat = all_r.ravel()
    j, t = _interp_weights(flat, r)
    rho = np.nan_to_num(rho2_uw, nan=0.0, posinf=0.0, neginf=0.0)
    v = W.ravel() * ((1.0 - t) * rho[j] + t * rho[j + 1])
    ir = np.clip(np.searchsorted(r, flat), 1, len(r) - 1)
    il = ir - 1
    wr = (flat - r[il]) / (r[ir] - r[il] + 1e-30)
    out = np.zeros(len(r))
    np.add.at(out, il, v * (1.0 - wr))
    np.add.at(out, ir, v * wr)
    return out


def alpha_beta(s: _Setup):
    """(alpha_dir, beta_dir, alpha_ex, beta_ex) such that the quadrature part
    of I_L is alpha^T K_L beta for every L (length-N float64 vectors)."""
    N, n1 = s.N, s.idx_limit
    if not s.oscillatory:
        return s.rho1_dir_w, s.rho2_dir_w, s.rho1_ex_w, s.rho2_ex_w
    if n1 < 2:
        z = np.zeros(N)
        return z, z, z, z
    r1 = s.r[:n1]
    nodes = _filon_nodes(r1, s.k_total, s.phase_increment, s.eta_total, s.CC_nodes)
    w1 = np.zeros(N)
    w1[:n1] = s.w[:n1]
    clean = lambda v: np.nan_to_num(v, nan=0.0, posinf=0.0, neginf=0.0)
    if nodes is None:
        # standard weighted dot products of the filon / filon_exchange branches:
        # rows on r[:N1], columns over the full r2 grid for both families
        # (oscillatory_integrals.py:2571, :2587-2600, :2676-2691, :2703-2705)
        ad = w1 * clean(s.rho1_dir_uw)
        ae = w1 * clean(s.rho1_ex_uw)
        be = s.w * clean(s.rho2_ex_uw)
    else:
        ad = filon_alpha(s.rho1_dir_uw, s.r, n1, nodes)
        ae = filon_alpha(s.rho1_ex_uw, s.r, n1, nodes)
        be = filon_exchange_beta(s.rho2_ex_uw, s.r, nodes)
    bd = s.w * clean(s.rho2_dir_uw)
    return ad, bd, ae, be


def radial_ME_all_L_reduced(r, w_simpson, w_trapz, V_core, U_i, bound_i, bound_f, cont_i, cont_f,
                            L_max, use_oscillatory_quadrature=True, oscillatory_method="advanced",
                            CC_nodes=5, phase_increment=1.5708, min_grid_fraction=0.10,
                            k_threshold=0.5, U_f_array=None, is_excitation=None,
                            return_parts=False):
    """Same numbers as radial_ME_all_L_faithful through I = alpha^T K_L beta
    (exact O(N) contraction) + tails + L=0 corrections."""
    s = _Setup(r, w_simpson, w_trapz, V_core, U_i, bound_i, bound_f, cont_i, cont_f, L_max,
               use_oscillatory_quadrature, oscillatory_method, CC_nodes, phase_increment,
               min_grid_fraction, k_threshold, U_f_array, is_excitation)
    ad, bd, ae, be = alpha_beta(s)
    cd = semisep.il_all_L(s.r, ad, bd, s.L_max)
    ce = semisep.il_all_L(s.r, ae, be, s.L_max)
    I_dir, I_ex = {}, {}
    tails = np.zeros(s.L_max + 1)
    noise = np.zeros(s.L_max + 1)
    for L in range(s.L_max + 1):
        t = 0.0
        if s.oscillatory and s.method != "legacy":
            t += s.direct_tail(L)
            if return_parts:
                noise[L] = s.direct_tail_noise(L)
        if s.method == "legacy":
            t += s.legacy_tail(L)
        tails[L] = t
        d = cd[L] + t
        e = ce[L]
        if L == 0:
            d += s.l0_correction(s.rho1_dir_w, s.rho2_dir_w)
            e += s.l0_correction(s.rho1_ex_w, s.rho2_ex_w)
        I_dir[L] = float(d) if np.isfinite(d) else 0.0
        I_ex[L] = float(e) if np.isfinite(e) else 0.0
    res = Result(I_dir, I_ex)
    if return_parts:
        return res, dict(setup=s, alpha_dir=ad, beta_dir=bd, alpha_ex=ae, beta_ex=be,
                         contraction_dir=cd, contraction_ex=ce, tails=tails, tail_noise=noise)
    return res


def plain_faithful(r, alpha_dir, beta_dir, alpha_ex, beta_ex, L_values):
    """Plain contraction the way radial_ME_all_L evaluates it when oscillatory
    quadrature is off (dwba_matrix_elements.py:1057-1080, :1113-1116,
    :1243-1245, :1288-1289): full N x N ratio / inv_gtr / log_ratio
    matrices, then per L one N^2 exp and two GEMV pairs.  Used as the CPU
    baseline of the kernel-only workload.  Returns {L: (I_dir, I_ex)}."""
    r = np.asarray(r, dtype=np.float64)
    r1c, r2r = r[:, None], r[None, :]
    rl = np.minimum(r1c, r2r)
    rg = np.maximum(r1c, r2r)
    ratio = rl / rg
    del rl
    inv_gtr = 1.0 / rg
    del rg
    log_ratio = np.log(np.minimum(ratio, 1.0 - 1e-12) + 1e-300)
    del ratio
    out = {}
    for L in L_values:
        K = inv_gtr if L == 0 else inv_gtr * np.exp(L * log_ratio)
        out[L] = (float(np.dot(alpha_dir, np.dot(K, beta_dir))),
                  float(np.dot(alpha_ex, np.dot(K, beta_ex))))
        del K
    return out
