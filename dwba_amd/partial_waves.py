"""Projectile partial-wave sums of a scan, batched on the GPU.

Replaces the per-pair loops around the integrals:

* excitation: the GPU branch of compute_total_excitation_cs
  (driver.py:877-1152) -- for every l_i, every l_f of the transition's
  parity up to max(L_max_proj, l_i + L_max_integrals), one integral call and
  one calculate_amplitude_contribution per (M_i, M_f), the per-l_i
  amplitudes added to the running total, convergence tested on the total
  after each l_i -- and the cross sections of section 6 (:1232-1260);
* ionization: the partial-wave loop of _compute_sdcs_at_energy
  (ionization.py:526-705) -- triples (l_i, l_ej, l_f), angle-free
  coefficients, coherent sums over l_i per (l_ej, l_f, M_i, M_f), the
  3-in-a-row relative-change stop.

Here every pair of every energy goes through ONE batched integral launch
(ScanBatch), the amplitudes are summed per (energy, l_i) slot on the device
(csrc/amplitude.hip), the running totals after each l_i and their DCS/TCS
are formed on the device, and only then are the loop's stopping rules
replayed on the host, l_i by l_i, on those few numbers: the reference stops
adding at the first l_i that meets a rule, and the totals kept are the
running total at that point -- so computing the later l_i as well changes no
result.  Waves are taken from the caller's caches (the reference's
precompute_continuum_waves / chi caches); a missing wave (None) is skipped
as the reference skips it.
"""
from __future__ import annotations

import math
from collections import deque
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence

import numpy as np
import torch

from . import _hip
from .amplitudes import AmplitudeAssembler, IonizationAssembler

A0_SQUARED_CM2 = 2.8002852e-17          # sigma_total.py:66


def sigma_au_to_cm2(x):
    """sigma_total.py:224-244."""
    return x * A0_SQUARED_CM2


# ------------------------------------------------------------------ excitation
@dataclass
class ExcitationEnergy:
    """One incident energy of an excitation scan: grid, potentials, bound
    orbitals, wave numbers and the continuum-wave caches (l -> wave or None),
    as compute_total_excitation_cs holds them before its partial-wave loop."""
    grid: object
    V_core: np.ndarray
    U_i: np.ndarray
    U_f: np.ndarray
    orb_i: object
    orb_f: object
    k_i: float
    k_f: float
    chi_i: Dict[int, object]
    chi_f: Dict[int, object]
    L_max_proj: int


@dataclass
class ExcitationChannel:
    """The fields of the reference's channel spec the loop reads."""
    L_target_i: int
    L_target_f: int
    L_max_integrals: int
    N_equiv: float = 1.0


@dataclass
class ExcitationResult:
    """Per energy: the totals the loop leaves (F, G per projection pair,
    complex, (n_combos, n_theta)), the summed DCS and its integral (a.u. and
    cm^2, before any top-up), the per-l_i contributions in cm^2 (keys
    "L{l_i}", as partial_waves_dict), and where and why the sum stopped."""
    combos: list
    F: np.ndarray
    G: np.ndarray
    dcs: np.ndarray
    sigma_au: float
    sigma_cm2: float
    partial_waves: Dict[str, float]
    stop_reason: Optional[str]
    stop_L: Optional[int]
    skipped_li: int
    n_pairs: int


def excitation_pairs(energies: Sequence[ExcitationEnergy], chan: ExcitationChannel):
    """The (energy, l_i, l_f) triples the loop visits, in its order
    (driver.py:877-1000): l_i ascending with its chi_i present, l_f from 0 to
    max(L_max_proj, l_i + L_max_integrals) with (l_i + l_f) % 2 ==
    (L_i + L_f) % 2 and its chi_f present.  Returns (triples, has_chi_i),
    has_chi_i[e][l_i] telling skipped l_i apart."""
    par = (chan.L_target_i + chan.L_target_f) % 2
    out, has = [], []
    for e, en in enumerate(energies):
        h = []
        for l_i in range(en.L_max_proj + 1):
            ok = en.chi_i.get(l_i) is not None
            h.append(ok)
            if not ok:
                continue
            for l_f in range(0, max(en.L_max_proj, l_i + chan.L_max_integrals) + 1):
                if (l_i + l_f) % 2 != par or en.chi_f.get(l_f) is None:
                    continue
                out.append((e, l_i, l_f))
        has.append(h)
    return out, has


def excitation_partial_wave_sums(energies: Sequence[ExcitationEnergy], chan: ExcitationChannel, theta_grid,
                                 oscillatory_method: str = "advanced", CC_nodes: int = 5,
                                 phase_increment: float = 1.5708, min_grid_fraction: float = 0.10,
                                 k_threshold: float = 0.5, precision: str = "fp64") -> List[ExcitationResult]:
    """The partial-wave loop of every energy, all pairs in one batched launch."""
    from .batch import EnergyInput, ScanBatch
    triples, has = excitation_pairs(energies, chan)
    if not triples:
        raise ValueError("no channel pairs: every wave is missing")
    inputs = []
    for e, en in enumerate(energies):
        prs = [(en.chi_i[li], en.chi_f[lf]) for (ee, li, lf) in triples if ee == e]
        inputs.append(EnergyInput(grid=en.grid, V_core=en.V_core, U_i=en.U_i, bound_i=en.orb_i,
                                  bound_f=en.orb_f, pairs=prs, U_f=en.U_f))
    keep = [i for i, en in enumerate(inputs) if en.pairs]
    b = ScanBatch([inputs[i] for i in keep], chan.L_max_integrals, precision=precision,
                  oscillatory_method=oscillatory_method, CC_nodes=CC_nodes, phase_increment=phase_increment,
                  min_grid_fraction=min_grid_fraction, k_threshold=k_threshold)
    b.run()
    # device row of each triple: ScanBatch stores pair k of its list at row out_index[k]
    rows = np.asarray(b.out_index, dtype=np.int64)
    return excitation_sums_from_integrals(b.out, rows, triples, has, energies, chan, theta_grid)


def excitation_sums_from_integrals(integrals: torch.Tensor, rows, triples, has_chi_i, energies, chan,
                                   theta_grid) -> List[ExcitationResult]:
    """The amplitude and convergence part of the loop on integrals already
    on the device: integrals[rows[q]] holds the (direct, exchange) I_L of
    triples[q] = (energy, l_i, l_f)."""
    n_e = len(energies)
    Li, Lf = chan.L_target_i, chan.L_target_f
    theta = np.ascontiguousarray(theta_grid, dtype=np.float64)
    asm = AmplitudeAssembler(theta, Li, Lf, chan.L_max_integrals, N_equiv=chan.N_equiv,
                             device=integrals.device)
    nc, nt = len(asm.combos), theta.size
    # slots: (energy, l_i) for l_i = 0..L_max_proj, energies back to back
    first = np.zeros(n_e + 1, dtype=np.int64)
    for e, en in enumerate(energies):
        first[e + 1] = first[e] + en.L_max_proj + 1
    ns = int(first[-1])
    if ns > 65535:
        raise ValueError("more than 65535 (energy, l_i) slots in one call")
    T = np.asarray(triples, dtype=np.int64).reshape(-1, 3)
    pairs = np.stack([np.asarray(rows, dtype=np.int64), T[:, 1], T[:, 2], first[T[:, 0]] + T[:, 1]], axis=1)
    k_i = np.concatenate([np.full(en.L_max_proj + 1, en.k_i) for en in energies])
    k_f = np.concatenate([np.full(en.L_max_proj + 1, en.k_f) for en in energies])
    res = asm.excitation(integrals, pairs, k_i, k_f, n_slots=ns)
    lib = _hip.load_library()
    sh = _hip.stream_handle()
    dev = integrals.device
    Fr = torch.view_as_real(res.F).contiguous()
    Gr = torch.view_as_real(res.G).contiguous()
    cF = torch.empty_like(Fr)
    cG = torch.empty_like(Gr)
    flags = torch.zeros(ns, dtype=torch.int32, device=dev)
    d_first = torch.as_tensor(first.astype(np.int32), device=dev)
    _hip.check(lib.dwba_pw_cumulate_f64(n_e, _hip.ptr(d_first), nc, nt, _hip.ptr(Fr), _hip.ptr(Gr), _hip.ptr(cF),
                                        _hip.ptr(cG), _hip.ptr(flags), sh), "dwba_pw_cumulate_f64")
    ok = (k_i > 0) & (k_f > 0)
    pref = np.where(ok, k_f / np.where(ok, k_i, 1.0), 0.0) * (chan.N_equiv / (2 * Li + 1))
    d_pref = torch.as_tensor(np.ascontiguousarray(pref), device=dev)
    snap = torch.empty((ns, nt), dtype=torch.float64, device=dev)
    snap_tcs = torch.empty((ns,), dtype=torch.float64, device=dev)
    _hip.check(lib.dwba_dcs_f64(ns, nc, nt, _hip.ptr(asm.theta), _hip.ptr(cF), _hip.ptr(cG), _hip.ptr(d_pref),
                                _hip.ptr(snap), _hip.ptr(snap_tcs), sh), "dwba_dcs_f64")
    snap_h = snap.cpu().numpy()
    snap_tcs_h = snap_tcs.cpu().numpy()
    slot_tcs_h = res.tcs.cpu().numpy()
    flags_h = flags.cpu().numpy()
    valid = np.zeros(ns, dtype=bool)
    valid[pairs[:, 3]] = True
    out = []
    for e, en in enumerate(energies):
        s0 = int(first[e])
        pw: Dict[str, float] = {}
        history: deque = deque(maxlen=4)
        stop_reason = stop_L = None
        last_added = -1
        skipped = 0
        for l_i in range(en.L_max_proj + 1):
            s = s0 + l_i
            if not has_chi_i[e][l_i]:
                skipped += 1
                continue
            if valid[s]:
                pw[f"L{l_i}"] = float(sigma_au_to_cm2(slot_tcs_h[s]))
            # numerical safety (driver.py:1073-1089): this l_i's amplitudes, then the total so far
            if flags_h[s] & 1:
                stop_reason, stop_L = "Numerical failure: NaN/Inf in l_i amplitudes", l_i
                break
            if last_added >= 0 and flags_h[last_added] & 2:
                stop_reason, stop_L = "Numerical failure: NaN/Inf in accumulated amplitudes", l_i
                break
            last_added = s
            # per-angle DCS stability of the total after adding l_i (:1108-1126)
            if l_i >= 5:
                history.append((l_i, snap_h[s]))
                if len(history) >= 4:
                    old = history[-4][1]
                    with np.errstate(divide="ignore", invalid="ignore"):
                        ch = np.abs(snap_h[s] - old) / (snap_h[s] + 1e-50)
                        ch[~np.isfinite(ch)] = 0.0
                    mx = float(np.max(ch))
                    if mx < 0.01 and l_i > 15:
                        stop_reason, stop_L = f"Converged: max Δ(DCS)/DCS = {mx:.2e} over all θ (stable)", l_i
                        break
        if last_added >= 0:
            F = torch.view_as_complex(cF[last_added]).cpu().numpy()
            G = torch.view_as_complex(cG[last_added]).cpu().numpy()
            dcs = snap_h[last_added].copy()
            sig = float(snap_tcs_h[last_added])
        else:
            F = np.zeros((nc, nt), complex)
            G = np.zeros((nc, nt), complex)
            dcs = np.zeros(nt)
            sig = 0.0
        out.append(ExcitationResult(asm.combos, F, G, dcs, sig, float(sigma_au_to_cm2(sig)), pw, stop_reason,
                                    stop_L, skipped, int(np.count_nonzero(T[:, 0] == e))))
    return out


# ------------------------------------------------------------------ ionization
@dataclass
class IonizationPoint:
    """One ejected-energy point of an ionization scan: the grid, the incident
    distorting potential, the ion's potential (V_core), the initial orbital,
    the wave numbers and the wave caches the loop reads (incident l_i,
    ejected l_ej, scattered l_f; None = failed solve)."""
    grid: object
    V_core: np.ndarray
    U_inc: np.ndarray
    orb_i: object
    k_i: float
    k_scatt: float
    k_eject: float
    chi_i: Dict[int, object]
    chi_eject: Dict[int, object]
    chi_scatt: Dict[int, object]
    L_max_proj: int
    l_f_max_scatt: int


@dataclass
class IonizationChannel:
    """The fields of IonizationChannelSpec the loop reads (ionization.py:135-150)."""
    l_i: int
    l_eject_max: int
    L_max: int
    N_equiv: float = 1.0
    convergence_threshold: float = 0.01


@dataclass
class IonizationResult:
    """Per point: sigma (a.u., before the high-L top-up), the per-l_i
    contributions in a.u. (keys "L{l_i}", as partial_L_contribs), the
    coherent coefficient sums per (l_ej, l_f) (combos x 2 complex: f, g) and
    the last l_i summed."""
    sigma_au: float
    partial_L: Dict[str, float]
    coherent: Dict[tuple, np.ndarray]
    combos: list
    last_l_i: int
    converged: bool
    n_pairs: int


def ionization_triples(points: Sequence[IonizationPoint], chan: IonizationChannel):
    """(point, l_i, l_ej, l_f) in the loop's order (ionization.py:526-580):
    l_i until the first missing incident wave, l_ej with its wave, l_f from 0
    to min(max(L_max_proj, l_i + L_max), l_f_max_scatt) of parity
    (l_i + l_f) % 2 == (l_i_target + l_ej) % 2 with its wave.  Returns
    (triples, n_li) with n_li[p] the number of l_i the loop reaches."""
    out, n_li = [], []
    for p, pt in enumerate(points):
        reach = 0
        for l_i in range(pt.L_max_proj + 1):
            if pt.chi_i.get(l_i) is None:
                break
            reach = l_i + 1
            for l_ej in range(chan.l_eject_max + 1):
                if pt.chi_eject.get(l_ej) is None:
                    continue
                lf_max = min(max(pt.L_max_proj, l_i + chan.L_max), pt.l_f_max_scatt)
                par = (chan.l_i + l_ej) % 2
                for l_f in range(0, lf_max + 1):
                    if (l_i + l_f) % 2 != par or pt.chi_scatt.get(l_f) is None:
                        continue
                    out.append((p, l_i, l_ej, l_f))
        n_li.append(reach)
    return out, n_li


def ionization_sdcs(points: Sequence[IonizationPoint], chan: IonizationChannel,
                    oscillatory_method: str = "advanced", CC_nodes: int = 5, phase_increment: float = 1.5708,
                    min_grid_fraction: float = 0.10, k_threshold: float = 0.5,
                    precision: str = "fp64") -> List[IonizationResult]:
    """The SDCS partial-wave loop of every point, all triples in one launch."""
    from .batch import EnergyInput, ScanBatch
    triples, n_li = ionization_triples(points, chan)
    if not triples:
        raise ValueError("no (l_i, l_ej, l_f) triples: every wave is missing")
    inputs = []
    for p, pt in enumerate(points):
        prs = [(pt.chi_i[li], pt.chi_scatt[lf], pt.chi_eject[le]) for (pp, li, le, lf) in triples if pp == p]
        if prs:
            inputs.append(EnergyInput(grid=pt.grid, V_core=pt.V_core, U_i=pt.U_inc, bound_i=pt.orb_i,
                                      bound_f=prs[0][2], pairs=prs, U_f=pt.V_core))
    b = ScanBatch(inputs, chan.L_max, precision=precision, oscillatory_method=oscillatory_method,
                  CC_nodes=CC_nodes, phase_increment=phase_increment, min_grid_fraction=min_grid_fraction,
                  k_threshold=k_threshold)
    b.run()
    rows = np.asarray(b.out_index, dtype=np.int64)
    return ionization_sums_from_integrals(b.out, rows, triples, n_li, points, chan)


def ionization_sums_from_integrals(integrals: torch.Tensor, rows, triples, n_li, points, chan) -> List[IonizationResult]:
    """The coefficient and convergence part of the loop on device-resident
    integrals: integrals[rows[q]] holds the I_L of triples[q] = (point, l_i, l_ej, l_f)."""
    dev = integrals.device
    asm = IonizationAssembler(chan.l_i, chan.l_eject_max, chan.L_max, include_eject_norm=True, device=dev)
    nc = len(asm.combos)
    T = np.asarray(triples, dtype=np.int64).reshape(-1, 4)
    n = T.shape[0]
    ki = np.array([points[p].k_i for p in T[:, 0]])
    kf = np.array([points[p].k_scatt for p in T[:, 0]])
    ke = np.array([points[p].k_eject for p in T[:, 0]])
    a, _ = asm._coeffs(integrals, np.stack([np.asarray(rows)[np.arange(n)], T[:, 1], T[:, 3], T[:, 2]], axis=1),
                       ki, kf, ke)
    # runs: pairs of one (point, l_ej, l_f) in ascending l_i (stable sort keeps the loop order within a key)
    key = T[:, 0] * 1_000_000 + T[:, 2] * 1000 + T[:, 3]
    order = np.lexsort((T[:, 1], key))
    ks = key[order]
    starts = np.flatnonzero(np.r_[True, ks[1:] != ks[:-1]])
    run_first = np.r_[starts, n].astype(np.int32)
    lib = _hip.load_library()
    sh = _hip.stream_handle()
    d_rf = torch.as_tensor(run_first, device=dev)
    d_rp = torch.as_tensor(order.astype(np.int32), device=dev)
    dcoh = torch.empty((n, nc), dtype=torch.float64, device=dev)
    inc = torch.empty((n, nc), dtype=torch.float64, device=dev)
    _hip.check(lib.dwba_ion_spin_terms_f64(len(starts), _hip.ptr(d_rf), _hip.ptr(d_rp), nc, _hip.ptr(a),
                                           _hip.ptr(dcoh), _hip.ptr(inc), sh), "dwba_ion_spin_terms_f64")
    # per-(point, l_i) slot sums in the loop order (l_ej, then l_f)
    first = np.zeros(len(points) + 1, dtype=np.int64)
    for p, pt in enumerate(points):
        first[p + 1] = first[p] + pt.L_max_proj + 1
    slot = first[T[:, 0]] + T[:, 1]
    ns = int(first[-1])
    so = np.argsort(slot, kind="stable")
    cnt = np.bincount(slot, minlength=ns)
    sf = np.r_[0, np.cumsum(cnt)].astype(np.int32)
    d_sf = torch.as_tensor(sf, device=dev)
    d_sp = torch.as_tensor(so.astype(np.int32), device=dev)
    coh = torch.empty(ns, dtype=torch.float64, device=dev)
    incs = torch.empty(ns, dtype=torch.float64, device=dev)
    _hip.check(lib.dwba_slot_sums_f64(ns, _hip.ptr(d_sf), _hip.ptr(d_sp), nc, _hip.ptr(dcoh), _hip.ptr(inc),
                                      _hip.ptr(coh), _hip.ptr(incs), sh), "dwba_slot_sums_f64")
    coh_h = coh.cpu().numpy()
    inc_h = incs.cpu().numpy()
    a_h = torch.view_as_complex(a.view(n, nc, 2, 2)).cpu().numpy()
    out = []
    for p, pt in enumerate(points):
        pref = (pt.k_scatt * pt.k_eject / pt.k_i) * (float(chan.N_equiv) / float(2 * chan.l_i + 1)) * (2 * math.pi) ** 4
        partial: Dict[str, float] = {}
        total, prev, small, last, conv = 0.0, 0.0, 0, -1, False
        for l_i in range(n_li[p]):
            s = int(first[p]) + l_i
            partial[f"L{l_i}"] = float(pref * inc_h[s])
            total += coh_h[s]
            last = l_i
            sig = pref * total
            if l_i > 3 and prev > 1e-40:
                if abs(sig - prev) / prev < chan.convergence_threshold:
                    small += 1
                    if small >= 3:
                        conv = True
                        break
                else:
                    small = 0
            prev = sig
        sel = (T[:, 0] == p) & (T[:, 1] <= last)
        coherent: Dict[tuple, np.ndarray] = {}
        for q in np.flatnonzero(sel):
            kk = (int(T[q, 2]), int(T[q, 3]))
            coherent[kk] = coherent.get(kk, 0) + a_h[q]
        out.append(IonizationResult(float(pref * total), partial, coherent, asm.combos, last, conv,
                                    int(np.count_nonzero(T[:, 0] == p))))
    return out
