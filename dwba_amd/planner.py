"""Host-side control decisions of the I_L path (integers and a few scalars).

These are the branch decisions radial_ME_all_L takes before its O(N^2) work
(dwba_matrix_elements.py:884-1108, oscillatory_integrals.py:2555-2628): the
split index idx_limit, the phase increment, the phase-interval layout, the
log-spacing tests, and the Clenshaw-Curtis / Levin reference tables.  All
arithmetic on grid-sized data (densities, alpha/beta, contraction, tails)
runs on the GPU; this module only produces the parameters those kernels
take.  Per-grid prefix statistics make every per-pair decision O(log N).
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field

import numpy as np

PI = math.pi


# ---------------------------------------------------------------- tables
def cc_reference(n: int):
    """Reference Clenshaw-Curtis nodes / weights on [-1, 1] for n nodes.

    n == 5: the module-level table of oscillatory_integrals.py:120-142, whose
    last cosine term enters with weight 3/2 (weights sum to 31/15).
    Otherwise: clenshaw_curtis_nodes(n, -1, 1) (oscillatory_integrals.py:1539-1618),
    exact for polynomials of degree <= n-2.  Nodes run from +1 down to -1.
    """
    if n < 2:
        raise ValueError("Clenshaw-Curtis requires n_nodes >= 2.")
    theta = np.pi * np.arange(n) / (n - 1)
    if n == 5:
        x = np.cos(theta)
        jm = (n - 1) // 2
        jv = np.arange(1, jm + 1)
        ws = np.sum(np.cos(2 * jv[:, None] * theta[None, :]) / (4 * jv ** 2 - 1)[:, None], axis=0)
        ws += 0.5 * np.cos(2 * jm * theta) / (4 * jm ** 2 - 1)
        return x, (2.0 / (n - 1)) * (1 - 2 * ws)
    if n == 2:
        return np.array([1.0, -1.0]), np.array([0.5, 0.5]) * 2.0
    x = 0.5 * 2.0 * (np.cos(theta) + 1) + -1.0
    m = n - 1
    jm = m // 2
    w = np.ones(n)
    if jm > 0:
        jv = np.arange(1, jm + 1)
        w += np.sum((2.0 / (1 - 4 * jv ** 2))[:, None] * np.cos(2 * jv[:, None] * theta[None, :]), axis=0)
    if m % 2 == 0:
        w += np.cos(m * theta) / (1 - m ** 2)
    w *= 2.0 / m
    w[0] /= 2
    w[-1] /= 2
    w *= 0.5 * 2.0
    return x, w


def levin_tables(n: int = 8):
    """Chebyshev differentiation matrix on the descending Lobatto points
    cos(pi i/(n-1)) (oscillatory_integrals.py:970-1017) and the ascending node
    positions the Levin segments use (:1145-1147)."""
    x = np.cos(np.pi * np.arange(n) / (n - 1))
    c = np.ones(n)
    c[0] = c[-1] = 2.0
    i = np.arange(n)[:, None]
    j = np.arange(n)[None, :]
    xd = x[:, None] - x[None, :]
    np.fill_diagonal(xd, 1.0)
    D = (c[:, None] / c[None, :]) * ((-1.0) ** (i + j)) / xd
    np.fill_diagonal(D, 0.0)
    D[np.diag_indices(n)] = -np.sum(D, axis=1)
    return np.ascontiguousarray(D), np.ascontiguousarray(np.cos(np.pi * np.arange(n) / (n - 1))[::-1])


def filon_limits():
    """DWBA_MAX_FILON_SEGMENTS (>= 128, default 4096) and DWBA_MAX_EFFECTIVE_DPHI
    (>= pi/16, default pi/2), read like oscillatory_integrals.py:54-80."""
    def _i(name, d, lo):
        raw = os.environ.get(name, "").strip()
        try:
            return max(lo, int(raw)) if raw else max(lo, int(d))
        except Exception:
            return max(lo, int(d))

    def _f(name, d, lo):
        raw = os.environ.get(name, "").strip()
        try:
            return max(lo, float(raw)) if raw else max(lo, float(d))
        except Exception:
            return max(lo, float(d))
    return _i("DWBA_MAX_FILON_SEGMENTS", 4096, 128), _f("DWBA_MAX_EFFECTIVE_DPHI", np.pi / 2, np.pi / 16)


def tail_sweep_nodes(k_i, l_i, eta_i, k_f, l_f, eta_f, r_m, r_max, max_segments=4096, max_dphi=np.pi / 2):
    """Nodes the outer tail of one pair sweeps over all L, per phase sign, as
    compute_outer_integral_oscillatory chooses them (oscillatory_integrals.py:
    1361-1451) with the sinA*sinB phases of :350-456: omega = |Phi'(mid)|;
    Filon when |Phi''(mid)| h^2 < 0.1 (h = (pi/4)/omega): n_seg segments of
    neff nodes (segment cap and node up-scaling of :1280-1354), of which a lane
    sweeps neff - 1 per segment (shared end nodes once) plus one; else Levin,
    8 nodes per segment; nothing when |Phi'(mid)| < 1e-10 (the per-L
    Gauss-Legendre branch) or r_max <= r_m.  Returns (filon_nodes, levin_nodes)."""
    filon = levin = 0
    if r_max <= r_m + 1e-10:
        return 0, 0
    mid = 0.5 * (r_m + r_max)
    for sgn in (-1.0, 1.0):
        d1 = (k_i - k_f) if sgn < 0 else (k_i + k_f)
        d2 = 0.0
        if abs(eta_i) > 1e-15:
            d1 += eta_i / mid
            d2 -= eta_i / (mid * mid)
        if abs(eta_f) > 1e-15:
            d1 += sgn * eta_f / mid
            d2 -= sgn * eta_f / (mid * mid)
        if l_i > 0:
            d1 += l_i * (l_i + 1) / (2.0 * k_i * mid * mid)
            d2 -= l_i * (l_i + 1) / (k_i * mid ** 3)
        if l_f > 0:
            d1 += sgn * l_f * (l_f + 1) / (2.0 * k_f * mid * mid)
            d2 -= sgn * l_f * (l_f + 1) / (k_f * mid ** 3)
        om = abs(d1)
        if om < 1e-10:
            continue
        h = (np.pi / 4) / om
        span = r_max - r_m
        if abs(d2) * h * h < 0.1:
            nseg = min(max(1, int(np.ceil(span / h))), int(max_segments))
            eff = om * span / nseg
            neff = 5
            if eff > max_dphi:
                neff = min(15, max(5, 5 + 2 * (int(np.ceil(eff / max_dphi)) - 1)))
            filon += nseg * (neff - 1) + 1
        else:
            levin += 8 * max(1, int(np.ceil(span / h)))
    return filon, levin


# ---------------------------------------------------------------- grid stats
@dataclass
class GridStats:
    """Prefix statistics of one grid, computed once per grid."""
    r: np.ndarray
    w_trapz: np.ndarray
    dr_prefix_max: np.ndarray          # max(dr[:k]) for k = 1..N-1  (index k-1)
    dlog_cummean: np.ndarray
    dlog_cummax: np.ndarray
    dlog_cummin: np.ndarray
    positive_increasing: bool
    _coulomb_cache: dict = field(default_factory=dict)

    @classmethod
    def of(cls, r, w_trapz):
        r = np.asarray(r, dtype=np.float64)
        dr = np.diff(r)
        ok = bool(len(r) >= 1 and r[0] > 0.0 and np.all(dr > 0.0))
        if ok and len(r) >= 2:
            dl = np.diff(np.log(r))
            cm = np.cumsum(dl) / np.arange(1, len(dl) + 1)
            mx = np.maximum.accumulate(dl)
            mn = np.minimum.accumulate(dl)
        else:
            cm = mx = mn = np.zeros(0)
        return cls(r=r, w_trapz=np.asarray(w_trapz, dtype=np.float64),
                   dr_prefix_max=np.maximum.accumulate(np.abs(dr)) if len(dr) else np.zeros(0),
                   dlog_cummean=cm, dlog_cummax=mx, dlog_cummin=mn, positive_increasing=ok)

    def log_spaced(self, n: int, rel_tol: float = 5e-3) -> bool:
        """_is_log_spaced_grid on r[:n] (oscillatory_integrals.py:237-255): every
        log step within rel_tol of the mean step.  The prefix mean is the
        running sum over the steps, so it can differ from numpy's pairwise
        mean in the last bits; decisions are exact away from the threshold."""
        if n < 3 or not self.positive_increasing:
            return False
        k = n - 2                         # number of steps - 1
        mean = self.dlog_cummean[k]
        if not np.isfinite(mean) or mean <= 0.0:
            return False
        dev = max(self.dlog_cummax[k] - mean, mean - self.dlog_cummin[k])
        return bool(dev <= rel_tol * mean)

    def max_phase_step(self, n: int, k_total: float, eta_total: float) -> float:
        """max over r[:n] of the phase change per step (check_phase_sampling,
        oscillatory_integrals.py:1456-1509)."""
        if n < 2:
            return 0.0
        if abs(eta_total) > 1e-10:
            key = (float(k_total), float(eta_total))
            pm = self._coulomb_cache.get(key)
            if pm is None:
                r = self.r
                r_mid = np.maximum(0.5 * (r[:-1] + r[1:]), 1e-10)
                pm = np.maximum.accumulate(np.abs((k_total + eta_total / r_mid) * np.diff(r)))
                if len(self._coulomb_cache) > 64:
                    self._coulomb_cache.clear()
                self._coulomb_cache[key] = pm
            return float(pm[n - 2])
        return float(abs(k_total) * self.dr_prefix_max[n - 2])


# ---------------------------------------------------------------- decisions
def _searchsorted(r, x) -> int:
    return int(np.searchsorted(r, x))


def split_index(gs: GridStats, U_i, U_f, u_i, u_f, k_i, k_f, l_i, l_f, idx_match_i, idx_match_f,
                min_grid_fraction=0.10, extent_cache=None) -> int:
    """idx_limit of radial_ME_all_L (dwba_matrix_elements.py:909-1022): the
    earlier match point, raised past the centrifugal turning point (at least
    min_grid_fraction of the grid), extended to the 99% extent of u_i and
    u_f (trapezoid-weighted cumulative probability), then pushed out (at most
    500 points) until max(|U_i|, |U_f|) + l(l+1)/(2r^2) < 0.03 k_min^2/2."""
    r = gs.r
    n = len(r)
    idx = n
    if 0 < idx_match_i < n:
        idx = min(idx, idx_match_i + 1)
    if 0 < idx_match_f < n:
        idx = min(idx, idx_match_f + 1)
    k_eff = min(k_i, k_f) if min(k_i, k_f) > 0.1 else max(k_i, k_f)
    lw = max(l_i, l_f)
    if k_eff > 1e-3:
        it = min(_searchsorted(r, (lw + 0.5) / k_eff), n - 20)
        floor_idx = min(max(it + 20, int(n * min_grid_fraction)), n)
    else:
        floor_idx = max(1, int(n * min_grid_fraction))
    idx = max(idx, floor_idx)
    ext = 0
    for u in (u_i, u_f):
        e = None
        if extent_cache is not None:
            e = extent_cache.get(id(u))
            if e is not None and e[0] is not u:
                e = None
        if e is None:
            u_sq = np.asarray(u, dtype=np.float64) ** 2
            val = 0
            if np.sum(u_sq) > 1e-30:
                pc = np.cumsum(u_sq * gs.w_trapz)
                pc /= (pc[-1] + 1e-300)
                val = int(np.searchsorted(pc, 0.99))
            e = (u, val)
            if extent_cache is not None:
                extent_cache[id(u)] = e
        ext = max(ext, e[1])
    if ext > idx:
        idx = ext + 1
    kin = 0.5 * min(k_i, k_f) ** 2
    if kin > 1e-10:
        cent = lw * (lw + 1)

        def veff(t):
            ui = abs(U_i[t])
            uf = abs(U_f[t]) if U_f is not None else 0.0
            return max(ui, uf) + cent / (2.0 * r[t] ** 2)

        if veff(max(0, idx - 1)) / kin > 0.03:
            # first t in [idx, idx+500) with V_eff(t)/E < 0.03 (vectorised scan)
            t1 = min(idx + 500, n)
            if t1 > idx:
                rr = r[idx:t1]
                ui = np.abs(np.asarray(U_i[idx:t1], dtype=np.float64))
                v = np.maximum(ui, np.abs(np.asarray(U_f[idx:t1], dtype=np.float64))) if U_f is not None \
                    else np.maximum(ui, 0.0)
                v = v + cent / (2.0 * rr ** 2)
                hit = np.nonzero(v / kin < 0.03)[0]
                if len(hit):
                    idx = idx + int(hit[0]) + 1
    return int(idx)


def phase_increment(gs: GridStats, n1: int, k_total: float, requested: float, eta_total: float,
                    threshold=PI / 4, min_increment=PI / 16) -> float:
    """_adaptive_phase_increment_for_grid (oscillatory_integrals.py:1716-1755)."""
    d_req = min(float(max(min_increment, requested)), threshold)
    if k_total < 1e-10 or n1 < 2:
        return d_req
    mp = gs.max_phase_step(n1, k_total, eta_total)
    if not (mp > threshold):
        return d_req
    scale = float(threshold) / max(mp, 1e-12)
    return float(np.clip(d_req * max(0.25, min(1.0, scale)), min_increment, threshold))


@dataclass
class FilonLayout:
    """Phase-interval layout of one pair's Filon branches."""
    mode: int                # DWBA_PREP_* value
    p_start: float = 0.0
    p_step: float = 0.0
    p_end: float = 0.0
    n_pnodes: int = 0
    coulomb: bool = False    # nodes generated by stepping on the device
    k_total: float = 0.0
    eta_total: float = 0.0


PREP_PLAIN, PREP_FILON, PREP_STANDARD, PREP_ZERO = 0, 1, 2, 3


def filon_layout(gs: GridStats, n1: int, k_total: float, phase_inc: float, eta_total: float) -> FilonLayout:
    """Branch structure of oscillatory_kernel_integral_2d's filon /
    filon_exchange methods (oscillatory_integrals.py:2441-2628)."""
    if n1 < 2:
        return FilonLayout(PREP_ZERO)
    dphi = phase_increment(gs, n1, k_total, phase_inc, eta_total)
    if k_total < 1e-6:
        return FilonLayout(PREP_STANDARD)
    a, b = float(gs.r[0]), float(gs.r[n1 - 1])
    if abs(eta_total) < 1e-10:
        dr = dphi / k_total
        n_pts = int(np.ceil((b - a) / dr)) + 1
        if n_pts - 1 < 1:
            return FilonLayout(PREP_STANDARD)
        step = (b - a) / (n_pts - 1)
        if not step > 1e-12:
            return FilonLayout(PREP_STANDARD)
        return FilonLayout(PREP_FILON, a, step, b, n_pts, False, k_total, eta_total)
    if not (b - a) > 1e-12:
        return FilonLayout(PREP_STANDARD)
    return FilonLayout(PREP_FILON, a, dphi, b, 0, True, k_total, eta_total)


def xcd_runs(units, n_xcd: int = 8, pad=None) -> list:
    """Order work for a kernel that deals its workgroups to the XCDs in
    consecutive runs (item q of n runs on XCD q // ceil(n / n_xcd)): units are
    (cost, [items]) that should share one XCD's L2 (one energy's slices);
    they are spread over n_xcd bins by longest-processing-time (heaviest
    unit to the lightest bin), each bin keeps its units heaviest first, and
    the bins are padded with `pad` to one length and concatenated."""
    bins = [[] for _ in range(n_xcd)]
    load = [0.0] * n_xcd
    for cost, items in sorted(units, key=lambda u: -u[0]):
        b = min(range(n_xcd), key=lambda i: load[i])
        bins[b].extend(items)
        load[b] += cost
    n = max((len(b) for b in bins), default=0)
    out = []
    for b in bins:
        out.extend(b + [pad] * (n - len(b)))
    return out


def filon_groups(plans) -> list:
    """Groups of FILON pairs whose node geometry is the same: one energy
    (same grid), split index N1, phase layout and interpolation kinds.
    plans: iterable of (energy index, PairPlan-like with mode, N1, log1, log2,
    layout).  Returns lists of positions in `plans`, in first-seen order;
    pairs in other modes or with N1 < 2 belong to no group.  The cell sums of
    the Filon alpha/beta depend only on these fields, not on the densities."""
    groups = {}
    for pos, (ei, plan) in enumerate(plans):
        if plan.mode != PREP_FILON or plan.N1 < 2:
            continue
        lay = plan.layout
        key = (ei, plan.N1, plan.log1, plan.log2, lay.p_start, lay.p_step, lay.p_end, lay.n_pnodes,
               bool(lay.coulomb), lay.k_total if lay.coulomb else 0.0, lay.eta_total if lay.coulomb else 0.0)
        groups.setdefault(key, []).append(pos)
    return list(groups.values())
