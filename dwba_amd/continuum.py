"""Continuum (distorted) waves on the GPU, many partial waves per call.

Drop-ins for solve_continuum_wave (continuum.py:1439-2200) and
precompute_continuum_waves (driver.py:486-540): chi_l(k, r) on the radial
grid with unit asymptotic amplitude, its phase shift and match point.

Every grid-sized step runs in csrc/continuum.hip: the propagation
(Johnson's log-derivative ratio recursion with the backward product, or
local-step Numerov, each with the reference's periodic rescaling), the
5-point derivatives, the tail least-squares sums, the asymptotic wave and
the stitching with its amplitude check, and the analytic Born bypass.  The
host makes the per-wave choices the reference makes from r, U, k and l
alone: bypass or not (match point at under 15% of the radial span), the
first propagated index (0.9 of the turning radius when the wave starts
inside the barrier), the match point, the start values, and then, from a
handful of numbers per wave, the phase shift (log-derivative or the tail fit).

Solvers: "johnson" tries Johnson and falls back to Numerov when its wave
vanishes at the match point, as the reference's chain does; "numerov" tries
Numerov first.  The reference's "auto" picks scipy's adaptive RK45 for low
l at moderate energies; no grid recursion reproduces an adaptive ODE
solver's rounding, so "auto" here takes the Johnson chain for every l (the
waves agree with RK45's to the ODE tolerance, not to rounding), and "rk45"
is refused.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence

import numpy as np
import torch
from scipy.special import loggamma, spherical_jn, spherical_yn

from . import _hip

HARTREE_TO_EV = 27.211386245988
EV_TO_HARTREE = 1.0 / HARTREE_TO_EV
JOHNSON, NUMEROV = 0, 1


@dataclass
class ContinuumWave:
    """continuum.py:68-123."""
    l: int
    k_au: float
    chi_of_r: np.ndarray
    phase_shift: float
    r_match: float = 0.0
    idx_match: int = -1
    eta: float = 0.0
    sigma_l: float = 0.0
    phase_method: str = "unknown"
    solver_method: str = "numerov"

    @property
    def u_of_r(self) -> np.ndarray:
        return self.chi_of_r


def k_from_E_eV(E_eV: float) -> float:
    """grid.py:105-132."""
    return float(np.sqrt(2.0 * max(E_eV * EV_TO_HARTREE, 0.0)))


def _match_point(r, U, k, l, threshold=1e-2, idx_start=0, z_ion=0.0, min_search_fraction=0.25):
    """_find_match_point (continuum.py:250-368): the first index in the search
    window past the turning-point and Coulomb radii where |2U| + l(l+1)/r^2 <
    threshold k^2, else the fallback index."""
    k2 = k * k
    N = len(r)
    frac = float(np.clip(min_search_fraction, 0.0, 0.9))
    s0 = max(idx_start + 50, int(frac * N))
    s1 = int(0.9 * N)
    r_turn = math.sqrt(l * (l + 1)) / k if k > 1e-6 else 0.0
    s0 = max(s0, int(np.searchsorted(r, r_turn)) + 10)
    r_safe = 2.5 * r_turn
    r_coul = 3.0 * max(l, abs(z_ion) / k) / k if (abs(z_ion) > 1e-6 and k > 1e-6) else 0.0
    if s0 < s1:
        rr = r[s0:s1]
        ok = (rr >= r_safe) & (rr >= r_coul) & (np.abs(2.0 * U[s0:s1]) + l * (l + 1) / (rr * rr) < threshold * k2)
        hit = np.flatnonzero(ok)
        if hit.size:
            i = s0 + int(hit[0])
            return i, float(r[i])
    it = int(np.searchsorted(r, r_safe)) if r_safe > 0 else 0
    ic = int(np.searchsorted(r, r_coul)) if r_coul > 0 else 0
    fb = min(max(s0, int(0.7 * N), it + 20, ic + 5), N - 10)
    return fb, float(r[fb])


def _riccati(l, rho):
    """(j^_l, j^_l', n^_l, n^_l') at rho (continuum.py:126-184)."""
    if rho < 1e-10:
        jh, jp = 0.0, 0.0
    else:
        jl = float(spherical_jn(l, rho))
        jh, jp = rho * jl, (math.cos(rho) if l == 0 else rho * float(spherical_jn(l - 1, rho)) - l * jl)
    if rho < 1e-10:
        return jh, jp, -1e30, -1e30
    yl = float(spherical_yn(l, rho))
    if not np.isfinite(yl) or abs(yl) > 1e50:
        return jh, jp, -1e30, -1e30
    if l == 0:
        return jh, jp, rho * yl, math.sin(rho)
    ym = float(spherical_yn(l - 1, rho))
    npr = -1e30 if (not np.isfinite(ym) or abs(ym) > 1e50) else rho * ym - l * yl
    return jh, jp, rho * yl, npr


def _coulomb_fg(l, eta, rho):
    """_coulomb_FG_asymptotic (continuum.py:187-248)."""
    if rho < 1e-10:
        return 0.0, 0.0, 1e30, 1e30
    sig = float(np.imag(loggamma(l + 1 + 1j * eta)))
    tb = rho + eta * math.log(2.0 * rho) - l * math.pi / 2.0 + sig
    ll1 = l * (l + 1)
    th = tb - (ll1 / (2.0 * rho) if rho > 1.0 else 0.0)
    amp = 1.0 + (ll1 - 2.0 * eta * eta) / (4.0 * rho * rho) if rho > 5.0 else 1.0
    tp = 1.0 + eta / rho + ll1 / (2.0 * rho * rho)
    return amp * math.sin(th), amp * tp * math.cos(th), amp * math.cos(th), -amp * tp * math.sin(th)


def _logderiv(Y, k, r_m, l, z_ion, chi_m):
    """Phase from the log-derivative at the match point (continuum.py:370-459)
    and the amplitude chi_m / (F cos d - G sin d) of :1986-2003."""
    rho = k * r_m
    Yr = Y / k
    if abs(z_ion) < 1e-3:
        jh, jp, nh, npr = _riccati(l, rho)
        num, den = Yr * jh - jp, Yr * nh - npr
        F, G = jh, nh
    else:
        F, Fp, G, Gp = _coulomb_fg(l, -z_ion / k, rho)
        num, den = Yr * F - Fp, Yr * G - Gp
    d = (math.pi / 2.0 if num > 0 else -math.pi / 2.0) if abs(den) < 1e-15 else math.atan(num / den)
    ref = F * math.cos(d) - G * math.sin(d)
    return d, (abs(chi_m) if abs(ref) < 1e-100 else chi_m / ref)


def _lsq(s):
    """Phase and amplitude of chi ~ a sin + b cos from the normal-equation sums."""
    ss, sc, cc, ys, yc = s[:5]
    det = ss * cc - sc * sc
    if det == 0.0 or not np.isfinite(det):
        return 0.0, 0.0
    a = (ys * cc - yc * sc) / det
    b = (yc * ss - ys * sc) / det
    A = float(np.hypot(a, b))
    if not np.isfinite(A) or A > 1e100:
        return 0.0, 0.0
    return A, float(np.arctan2(b, a))


def _renorm_interval(E_eV):
    """adaptive_renorm_interval (continuum.py:1562-1590)."""
    return 50 if E_eV > 5000.0 else (75 if E_eV > 1000.0 else 100)


def _grid_arrays(grid, U_channel):
    r = np.ascontiguousarray(grid.r, dtype=np.float64)
    w = np.ascontiguousarray(grid.w_trapz if hasattr(grid, "w_trapz") else np.gradient(r), dtype=np.float64)
    U = np.ascontiguousarray(getattr(U_channel, "U_of_r", U_channel), dtype=np.float64)
    if U.shape != r.shape:
        raise ValueError("solve_continuum_wave: U_channel/grid shape mismatch.")
    if r.size < 40:
        raise ValueError("grid too short")
    return r, w, U


def solve_continuum_waves(grid, U_channel, ls: Sequence[int], E_eV: float, z_ion: float = 0.0,
                          phase_extraction_method: str = "hybrid", solver: str = "johnson",
                          allow_analytic_bypass: bool = True) -> List[Optional[ContinuumWave]]:
    """solve_continuum_wave for every l of `ls` on one grid, potential and
    energy; None where the reference returns None (no usable wave)."""
    if E_eV <= 0.0:
        raise ValueError("solve_continuum_wave: channel energy must be >0 eV (open channel).")
    sv = solver.lower()
    if sv in ("auto", "johnson"):
        order = (JOHNSON, NUMEROV)
    elif sv == "numerov":
        order = (NUMEROV, JOHNSON)
    else:
        raise ValueError(f"solver {solver!r}: only the grid recursions 'johnson', 'numerov' (and 'auto', which "
                         f"takes the Johnson chain) are provided")
    pm = phase_extraction_method.lower()
    r, w, U = _grid_arrays(grid, U_channel)
    k = k_from_E_eV(E_eV)
    if k <= 0.0 or not np.isfinite(k):
        raise ValueError("solve_continuum_wave: invalid k from given E_eV.")
    N = r.size
    ls = [int(l) for l in ls]
    dev = torch.device("cuda", _hip.ensure_init())
    lib = _hip.load_library()
    sh = _hip.stream_handle()
    r_d = torch.as_tensor(r, device=dev)
    w_d = torch.as_tensor(w, device=dev)
    U_d = torch.as_tensor(U, device=dev)
    ld = N
    span = max(float(r[-1] - r[0]), 1e-30)
    out: List[Optional[ContinuumWave]] = [None] * len(ls)

    # analytic bypass (neutral channels whose potential is negligible early)
    byp, prop = [], []
    for q, l in enumerate(ls):
        _, r_chk = _match_point(r, U, k, l, threshold=1e-4, z_ion=z_ion, min_search_fraction=0.0)
        if allow_analytic_bypass and (r_chk - float(r[0])) / span < 0.15 and abs(z_ion) <= 1e-6:
            byp.append(q)
        else:
            prop.append(q)
    if byp:
        arr = (_hip.CwBypass * len(byp))()
        for i, q in enumerate(byp):
            arr[i].r, arr[i].w_trapz, arr[i].U = _hip.ptr(r_d), _hip.ptr(w_d), _hip.ptr(U_d)
            arr[i].k, arr[i].n, arr[i].l = k, N, ls[q]
        d_b = _hip.struct_array_to_device(arr, dev)
        chi_b = torch.empty((len(byp), ld), dtype=torch.float64, device=dev)
        del_b = torch.empty(len(byp), dtype=torch.float64, device=dev)
        _hip.check(lib.dwba_cw_bypass_f64(len(byp), _hip.ptr(d_b), ld, _hip.ptr(chi_b), _hip.ptr(del_b), sh),
                   "dwba_cw_bypass_f64")
        chi_h, del_h = chi_b.cpu().numpy(), del_b.cpu().numpy()
        for i, q in enumerate(byp):
            out[q] = ContinuumWave(ls[q], k, chi_h[i].copy(), float(del_h[i]), eta=0.0, sigma_l=0.0,
                                   phase_method="born", solver_method="analytic_bypass")
    if not prop:
        return out

    # per-wave set-up: first propagated index, match point, start values
    info = []
    for q in prop:
        l = ls[q]
        ell = float(l)
        start = 0
        if ell * (ell + 1.0) / (r[0] * r[0]) + 2.0 * U[0] - k * k > 0:
            rt = math.sqrt(l * (l + 1)) / k if k > 1e-6 else r[-1]
            f = int(np.searchsorted(r, rt * 0.9))
            if 0 < f < N - 20:
                start = f
        im, rm = _match_point(r, U, k, l, threshold=1e-4, idx_start=start, z_ion=z_ion)
        r0, r1 = r[start], r[start + 1]
        Q0 = ell * (ell + 1.0) / (r0 * r0) + 2.0 * U[start] - k * k
        # Johnson R_1 (continuum.py:1243-1269)
        if Q0 > 0:
            R1 = math.exp(-math.sqrt(Q0) * (r1 - r0))
        elif k * r0 > 1e-10 and k * r1 > 1e-10:
            c0, c1 = r0 * spherical_jn(l, k * r0), r1 * spherical_jn(l, k * r1)
            R1 = float(c0 / c1) if abs(c1) > 1e-100 else (r0 / r1) ** (l + 1)
        else:
            R1 = (r0 / r1) ** (l + 1)
        R1 = min(max(R1, -1e10), 1e10)
        # Numerov start (continuum.py:1843-1898)
        if Q0 > 0:
            n0, n1 = 1e-20, 1e-20 * math.exp(math.sqrt(Q0) * (r1 - r0))
        else:
            ks = ell * (ell + 1.0) / (r0 * r0) - 2.0 * z_ion / r0 - k * k if abs(z_ion) >= 1e-3 else -1.0
            if ks > 0:
                n0, n1 = 1e-20, 1e-20 * math.exp(math.sqrt(ks) * (r1 - r0))
            else:
                n0, n1 = r0 * float(spherical_jn(l, k * r0)), r1 * float(spherical_jn(l, k * r1))
                nv = max(abs(n0), abs(n1))
                if 0 < nv < 1e-10:
                    sc = 1e-10 / nv
                    n0, n1 = n0 * sc, n1 * sc
        info.append(dict(q=q, l=l, start=start, im=im, rm=rm, R1=R1, n0=n0, n1=n1))

    nw = len(info)
    chi = torch.zeros((nw, ld), dtype=torch.float64, device=dev)
    dchi = torch.zeros_like(chi)
    method = [None] * nw
    pending = list(range(nw))
    for meth in order:
        if not pending:
            break
        arr = (_hip.CwWave * len(pending))()
        for i, t in enumerate(pending):
            f = info[t]
            a = arr[i]
            a.r, a.U, a.k, a.r1_ratio, a.chi0, a.chi1 = _hip.ptr(r_d), _hip.ptr(U_d), k, f["R1"], f["n0"], f["n1"]
            a.n, a.l, a.start, a.method, a.renorm = N, f["l"], f["start"], meth, _renorm_interval(E_eV)
        d_w = _hip.struct_array_to_device(arr, dev)
        nb = int(lib.dwba_cw_workspace_bytes(len(pending), ld))
        ws = torch.empty(nb, dtype=torch.uint8, device=dev)
        c = torch.empty((len(pending), ld), dtype=torch.float64, device=dev)
        dc = torch.empty_like(c)
        st = torch.empty(len(pending), dtype=torch.int32, device=dev)
        _hip.check(lib.dwba_cw_propagate_f64(len(pending), _hip.ptr(d_w), ld, _hip.ptr(c), _hip.ptr(dc), _hip.ptr(ws),
                                             nb, _hip.ptr(st), sh), "dwba_cw_propagate_f64")
        if int(st.max()) != 0:
            raise RuntimeError("continuum wave rescaled more often than the event buffer holds")
        ims = torch.as_tensor([info[t]["im"] for t in pending], device=dev)
        cm = c[torch.arange(len(pending), device=dev), ims].cpu().numpy()
        still = []
        for i, t in enumerate(pending):
            if abs(cm[i]) < 1e-100:
                still.append(t)
            else:
                chi[t] = c[i]
                dchi[t] = dc[i]
                method[t] = meth
        pending = still

    # phase shifts
    ok = [t for t in range(nw) if method[t] is not None]
    if not ok:
        return out
    eta = -z_ion / k
    coul = abs(z_ion) >= 1e-3
    sig = {t: float(np.imag(loggamma(info[t]["l"] + 1 + 1j * eta))) for t in ok} if coul else {t: 0.0 for t in ok}
    fin = (_hip.CwFinish * nw)()
    for t in range(nw):
        f = info[t]
        nt = int(N * 0.15)
        f["tail"] = max(f["im"], N - nt)
        a = fin[t]
        a.r, a.w_trapz, a.k, a.delta = _hip.ptr(r_d), _hip.ptr(w_d), k, 0.0
        a.eta, a.sigma = (eta, sig.get(t, 0.0)) if coul else (0.0, 0.0)
        a.n, a.l, a.idx_match, a.tail, a.coulomb = N, f["l"], f["im"], f["tail"], int(coul)
    d_fin = _hip.struct_array_to_device(fin, dev)
    sums = torch.empty((nw, 6), dtype=torch.float64, device=dev)
    _hip.check(lib.dwba_cw_tail_sums_f64(nw, _hip.ptr(d_fin), ld, _hip.ptr(chi), _hip.ptr(sums), sh),
               "dwba_cw_tail_sums_f64")
    S = sums.cpu().numpy()
    ims = torch.as_tensor([info[t]["im"] for t in range(nw)], device=dev)
    ar = torch.arange(nw, device=dev)
    cmv = chi[ar, ims].cpu().numpy()
    dmv = dchi[ar, ims].cpu().numpy()
    deltas = np.zeros(nw)
    good = np.zeros(nw, dtype=bool)
    pmeth = [""] * nw
    for t in ok:
        f = info[t]
        l, im, rm = f["l"], f["im"], f["rm"]
        cm_, dm_ = float(cmv[t]), float(dmv[t])
        if pm == "logderiv":
            d, A = _logderiv(dm_ / cm_, k, rm, l, z_ion, cm_)
            tag = "logderiv"
        else:
            valid = N - f["tail"] >= 10 and S[t, 5] > 1e-100
            A, d = _lsq(S[t]) if (valid or pm == "lsq") else (0.0, 0.0)
            valid = valid and np.isfinite(d) and np.isfinite(A) and A > 1e-100
            tag = "lsq"
            if pm != "lsq" and not valid:
                d, A = _logderiv(dm_ / cm_, k, rm, l, z_ion, cm_)
                tag = "logderiv"
            elif pm != "lsq" and l <= 10:
                dl, _ = _logderiv(dm_ / cm_, k, rm, l, z_ion, cm_)
                dd = abs(dl - d)
                if min(dd, 2.0 * math.pi - dd) < 0.1:
                    tag = "lsq_validated"
        if not (np.isfinite(A) and np.isfinite(d) and abs(A) >= 1e-100):
            continue
        deltas[t] = d
        good[t] = True
        pmeth[t] = tag
        fin[t].delta = d
    d_fin = _hip.struct_array_to_device(fin, dev)
    scratch = torch.empty_like(chi)
    final = torch.empty_like(chi)
    _hip.check(lib.dwba_cw_stitch_f64(nw, _hip.ptr(d_fin), ld, _hip.ptr(chi), _hip.ptr(dchi), _hip.ptr(scratch),
                                      _hip.ptr(final), sh), "dwba_cw_stitch_f64")
    fh = final.cpu().numpy()
    names = {JOHNSON: "Johnson", NUMEROV: "Numerov"}
    for t in ok:
        if not good[t]:
            continue
        f = info[t]
        m = method[t]
        sname = names[m] + ("" if m == order[0] else " (fallback)")
        out[f["q"]] = ContinuumWave(f["l"], k, fh[t].copy(), float(deltas[t]), f["rm"], f["im"],
                                    eta if k > 0 else 0.0, sig.get(t, 0.0) if abs(eta) > 1e-10 else 0.0,
                                    phase_method=pm, solver_method=f"{sname}/{pmeth[t]}")
    return out


def solve_continuum_wave(grid, U_channel, l: int, E_eV: float, z_ion: float = 0.0, tail_fraction: float = 0.1,
                         rtol: float = 1e-6, atol: float = 1e-8, phase_extraction_method: str = "hybrid",
                         solver: str = "auto", allow_analytic_bypass: bool = True) -> Optional[ContinuumWave]:
    """Drop-in for continuum.solve_continuum_wave (continuum.py:1439): one l.
    tail_fraction, rtol and atol belong to the RK45 path and are accepted for
    the signature."""
    return solve_continuum_waves(grid, U_channel, [l], E_eV, z_ion, phase_extraction_method, solver,
                                 allow_analytic_bypass)[0]


def precompute_continuum_waves(L_max: int, E_eV: float, z_ion: float, U, grid, phase_extraction_method="hybrid",
                               solver="auto", allow_analytic_bypass=True) -> Dict[int, ContinuumWave]:
    """driver.py:486-540: waves l = 0..L_max of one energy, all in one batched
    solve (the reference runs one process per l); failed waves are left out."""
    ws = solve_continuum_waves(grid, U, list(range(int(L_max) + 1)), E_eV, z_ion, phase_extraction_method, solver,
                               allow_analytic_bypass)
    return {l: w for l, w in enumerate(ws) if w is not None}
