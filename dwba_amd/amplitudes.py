"""Scattering amplitudes and cross sections from the radial integrals (GPU).

Drop-ins for the host code downstream of the integrals:

  calculate_amplitude_contribution    dwba_coupling.py:285-504
  calculate_ionization_coefficients   dwba_coupling.py:507-634
  dcs_dwba, integrate_dcs_over_angles sigma_total.py:147-217

and the batched form of the driver's accumulation (driver.py:1041-1085) and
of the ionization SDCS amplitude sums (ionization.py:611-642):
``AmplitudeAssembler`` reads the integrals where the batched scan leaves them
(device rows [row][2][L_max+1], direct then exchange) and returns F, G,
DCS and TCS per slot (a slot is any group of pairs summed coherently: an
energy, or an (energy, l_i) for partial-wave contributions) without the
integrals leaving the device.

The angular coefficients Cd_L, Cx_L and the constant phases depend only on
angular momenta; they are formed once per channel type on the host, exactly
(dwba_amd.angular), and uploaded.  The sums over L, the Y_lm table on the
angle grid, the pair sums and the cross sections run in csrc/amplitude.hip.
There is no CPU path: without the library or a GPU every call raises.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _hip, angular


@dataclass
class Amplitudes:
    """f(theta), g(theta) of one projection pair (dwba_coupling.py:244-247)."""
    f_theta: np.ndarray
    g_theta: np.ndarray


@dataclass(frozen=True)
class IonizationAmplitudeCoeffs:
    """Angle-free ionization coefficients (dwba_coupling.py:249-264)."""
    f_coeff: complex
    g_coeff: complex


@dataclass
class AmplitudeResult:
    """Per slot and projection pair (combos[c] = (M_i, M_f)):
    F, G (n_slots, n_combos, n_theta) complex128; dcs (n_slots, n_theta) and
    tcs (n_slots,) float64 in atomic units, summed over the projection pairs
    (or None when not requested).  All on the device."""
    combos: list
    F: torch.Tensor
    G: torch.Tensor
    dcs: torch.Tensor | None
    tcs: torch.Tensor | None


class _TypeTable:
    """Coefficient and phase tables of the channel types seen so far, for one
    list of projection pairs; uploaded again when a type is added."""

    def __init__(self, combos, L_max, build):
        self.combos = list(combos)
        self.L_max = int(L_max)
        self.build = build                    # key -> rows of angular.*_channel
        self.index: dict = {}
        self.coef: list = []
        self.phase: list = []
        self.dev = None
        self.d_coef = self.d_phase = None

    def ids(self, keys, device):
        new = False
        out = np.empty(len(keys), dtype=np.int32)
        for q, k in enumerate(keys):
            t = self.index.get(k)
            if t is None:
                rows = self.build(k)
                c = np.zeros((len(self.combos), 2, self.L_max + 1))
                ph = np.zeros((len(self.combos), 4))
                for i, (cd, cx, pd, px) in enumerate(rows):
                    c[i, 0], c[i, 1] = cd, cx
                    ph[i] = (pd.real, pd.imag, px.real, px.imag)
                t = self.index[k] = len(self.coef)
                self.coef.append(c)
                self.phase.append(ph)
                new = True
            out[q] = t
        if new or self.dev != device:
            self.d_coef = torch.as_tensor(np.stack(self.coef), device=device).contiguous()
            self.d_phase = torch.as_tensor(np.stack(self.phase), device=device).contiguous()
            self.dev = device
        return out


_PAIR_DTYPE = np.dtype([("row", "<i4"), ("type", "<i4"), ("l_f", "<i4"), ("pad_", "<i4"), ("scale", "<f8")])


def _pairs_device(rows, types, l_f, scale, device):
    """dwba_amp_pair records (the struct's layout: four int32 and a double) on the device."""
    a = np.zeros(len(rows), dtype=_PAIR_DTYPE)
    a["row"], a["type"], a["l_f"], a["scale"] = rows, types, l_f, scale
    return torch.from_numpy(a.view(np.uint8).copy()).to(device)


def _type_ids(table, keys2, device):
    """Channel-type indices of the (a, b) key rows, building tables only for the distinct keys."""
    k = np.asarray(keys2, dtype=np.int64)
    uniq, inv = np.unique(k, axis=0, return_inverse=True)
    ids = table.ids([tuple(int(x) for x in u) for u in uniq], device)
    return ids[np.asarray(inv).reshape(-1)]


def _check_integrals(integrals, L_max):
    if not (isinstance(integrals, torch.Tensor) and integrals.is_cuda and integrals.dtype == torch.float64
            and integrals.is_contiguous()):
        raise ValueError("integrals must be a contiguous float64 CUDA tensor")
    if integrals.dim() != 3 or integrals.shape[1] != 2 or integrals.shape[2] != L_max + 1:
        raise ValueError(f"integrals must have shape (rows, 2, {L_max + 1}), got {tuple(integrals.shape)}")


class AmplitudeAssembler:
    """F, G, DCS and TCS of a target transition L_i -> L_f on one angle grid.

    combos: the projection pairs (M_i, M_f); default every M_i in -L_i..L_i and
    M_f in -L_f..L_f in the driver's loop order (driver.py:1039-1040).
    L_max: the multipole count of the integrals (rows of L_max+1 values).
    N_equiv: the channel's equivalent-electron factor in dcs_dwba."""

    def __init__(self, theta, L_target_i: int, L_target_f: int, L_max: int, N_equiv: float = 1.0, combos=None,
                 device=None):
        self.Li, self.Lf, self.L_max = int(L_target_i), int(L_target_f), int(L_max)
        if not 0 <= self.L_max <= 63:
            raise ValueError("L_max must be in [0, 63]")
        if combos is None:
            combos = [(mi, mf) for mi in range(-self.Li, self.Li + 1) for mf in range(-self.Lf, self.Lf + 1)]
        self.combos = [(int(a), int(b)) for a, b in combos]
        if not 1 <= len(self.combos) <= 64:
            raise ValueError("between 1 and 64 projection pairs")
        self.N_equiv = float(N_equiv)
        self.device = torch.device("cuda", _hip.ensure_init(None if device is None else torch.device(device).index))
        self.theta_h = np.ascontiguousarray(theta, dtype=np.float64)
        if self.theta_h.ndim != 1 or self.theta_h.size < 1:
            raise ValueError("theta must be a non-empty 1-d grid")
        self.theta = torch.as_tensor(self.theta_h, device=self.device)
        self.m_dir = (ctypes.c_int32 * len(self.combos))(*[mi - mf for mi, mf in self.combos])
        self.m_ex = (ctypes.c_int32 * len(self.combos))(*[mf - mi for mi, mf in self.combos])
        self.m_max = max(abs(mi - mf) for mi, mf in self.combos)
        self._ylm_l = -1
        self._ylm = None
        Li, Lf, cmb, Lm = self.Li, self.Lf, self.combos, self.L_max
        self.types = _TypeTable(cmb, Lm, lambda k: angular.excitation_channel(k[0], k[1], Li, Lf, cmb, Lm))

    def ylm(self, l_max: int) -> torch.Tensor:
        """Y_{l,m}(theta_j, 0) for l <= l_max, |m| <= m_max ((l, m + m_max, j) layout)."""
        if l_max > self._ylm_l:
            lib = _hip.load_library()
            nm = 2 * self.m_max + 1
            y = torch.empty(((l_max + 1) * nm * self.theta.numel(),), dtype=torch.float64, device=self.device)
            _hip.check(lib.dwba_ylm_f64(l_max, self.m_max, self.theta.numel(), _hip.ptr(self.theta), _hip.ptr(y),
                                        _hip.stream_handle()), "dwba_ylm_f64")
            self._ylm, self._ylm_l = y, l_max
        return self._ylm

    def excitation(self, integrals: torch.Tensor, pairs, k_i, k_f, n_slots: int | None = None,
                   cross_sections: bool = True) -> AmplitudeResult:
        """pairs: rows (row, l_i, l_f, slot) -- the integrals' row of the
        channel pair, its partial waves and the slot it is summed into, pairs
        of a slot summed in the order given; k_i, k_f: per slot (a.u.)."""
        _check_integrals(integrals, self.L_max)
        P = np.asarray(pairs, dtype=np.int64).reshape(-1, 4)
        if P.shape[0] == 0:
            raise ValueError("no pairs")
        rows, li, lf, slot = P[:, 0], P[:, 1], P[:, 2], P[:, 3]
        if rows.min() < 0 or rows.max() >= integrals.shape[0]:
            raise ValueError("pair row outside the integrals")
        if li.min() < 0 or lf.min() < 0:
            raise ValueError("negative partial wave")
        ns = int(slot.max()) + 1 if n_slots is None else int(n_slots)
        if slot.min() < 0 or slot.max() >= ns or ns > 65535:
            raise ValueError("slot outside [0, n_slots), n_slots <= 65535")
        k_i = np.broadcast_to(np.asarray(k_i, dtype=np.float64), (ns,))
        k_f = np.broadcast_to(np.asarray(k_f, dtype=np.float64), (ns,))
        order = np.argsort(slot, kind="stable")
        rows, li, lf, slot = rows[order], li[order], lf[order], slot[order]
        types = _type_ids(self.types, np.stack([li, lf], axis=1), self.device)
        scale = (2.0 / math.pi) / (k_i[slot] * k_f[slot])
        first = np.zeros(ns + 1, dtype=np.int32)
        np.add.at(first, slot + 1, 1)
        first = np.cumsum(first).astype(np.int32)
        return self._run(integrals, rows, types, lf, scale, first, ns, k_i, k_f, int(lf.max()), cross_sections)

    def _run(self, integrals, rows, types, lf, scale, first, ns, k_i, k_f, l_top, cross_sections):
        lib = _hip.load_library()
        sh = _hip.stream_handle()
        dev = self.device
        n, nc, nt = len(rows), len(self.combos), self.theta.numel()
        d_pairs = _pairs_device(rows, types, lf, scale, dev)
        d_first = torch.as_tensor(first, device=dev)
        a = torch.empty((n, nc, 4), dtype=torch.float64, device=dev)
        _hip.check(lib.dwba_amp_coeffs_f64(n, _hip.ptr(d_pairs), nc, self.L_max, _hip.ptr(integrals),
                                           _hip.ptr(self.types.d_coef), _hip.ptr(self.types.d_phase), _hip.ptr(a),
                                           sh), "dwba_amp_coeffs_f64")
        y = self.ylm(l_top)
        F = torch.empty((ns, nc, nt, 2), dtype=torch.float64, device=dev)
        G = torch.empty_like(F)
        _hip.check(lib.dwba_amp_sum_f64(ns, _hip.ptr(d_first), _hip.ptr(d_pairs), nc, self.m_dir, self.m_ex,
                                        self._ylm_l, self.m_max, nt, _hip.ptr(y), _hip.ptr(a), _hip.ptr(F),
                                        _hip.ptr(G), sh), "dwba_amp_sum_f64")
        dcs = tcs = None
        if cross_sections:
            ok = (k_i > 0) & (k_f > 0)
            pref = np.where(ok, k_f / np.where(ok, k_i, 1.0), 0.0) * (self.N_equiv / (2 * self.Li + 1))
            d_pref = torch.as_tensor(np.ascontiguousarray(pref), device=dev)
            dcs = torch.empty((ns, nt), dtype=torch.float64, device=dev)
            tcs = torch.empty((ns,), dtype=torch.float64, device=dev)
            _hip.check(lib.dwba_dcs_f64(ns, nc, nt, _hip.ptr(self.theta), _hip.ptr(F), _hip.ptr(G), _hip.ptr(d_pref),
                                        _hip.ptr(dcs), _hip.ptr(tcs), sh), "dwba_dcs_f64")
        return AmplitudeResult(self.combos, torch.view_as_complex(F), torch.view_as_complex(G), dcs, tcs)


class IonizationAssembler:
    """Batched calculate_ionization_coefficients (dwba_coupling.py:507-634)
    and the coherent SDCS sums over pairs (ionization.py:631-642).

    combos: every (M_i, M_f) with M_i in -L_i..L_i and M_f in -l_ej_max..l_ej_max;
    a pair with l_ej < |M_f| gets zero coefficients for that projection pair."""

    def __init__(self, L_target_i: int, l_eject_max: int, L_max: int, include_eject_norm: bool = True, device=None):
        self.Li, self.l_ej_max, self.L_max = int(L_target_i), int(l_eject_max), int(L_max)
        if not 0 <= self.L_max <= 63:
            raise ValueError("L_max must be in [0, 63]")
        self.combos = [(mi, mf) for mi in range(-self.Li, self.Li + 1)
                       for mf in range(-self.l_ej_max, self.l_ej_max + 1)]
        if len(self.combos) > 4096:
            raise ValueError("too many projection pairs")
        self.norm = bool(include_eject_norm)
        self.device = torch.device("cuda", _hip.ensure_init(None if device is None else torch.device(device).index))
        Li, cmb, Lm = self.Li, self.combos, self.L_max

        def build(k):
            l_i, l_f, l_ej = k
            ok = [i for i, (_, mf) in enumerate(cmb) if abs(mf) <= l_ej]
            sub = angular.ionization_channel(l_i, l_f, l_ej, Li, [cmb[i] for i in ok], Lm)
            rows = [(np.zeros(Lm + 1), np.zeros(Lm + 1), 0j, 0j)] * len(cmb)
            for i, r in zip(ok, sub):
                rows[i] = r
            return rows
        self.types = _TypeTable(cmb, Lm, build)

    def coefficients(self, integrals: torch.Tensor, pairs, k_i, k_f, k_ej) -> torch.Tensor:
        """pairs: rows (row, l_i, l_f, l_ej); k_i, k_f, k_ej per pair (a.u.).
        Returns (n_pairs, n_combos, 2) complex128 on the device: [..., 0] the
        f coefficient, [..., 1] the g coefficient."""
        a, _ = self._coeffs(integrals, pairs, k_i, k_f, k_ej)
        return torch.view_as_complex(a.view(a.shape[0], a.shape[1], 2, 2))

    def _coeffs(self, integrals, pairs, k_i, k_f, k_ej):
        _check_integrals(integrals, self.L_max)
        P = np.asarray(pairs, dtype=np.int64).reshape(-1, 4)
        n = P.shape[0]
        if n == 0:
            raise ValueError("no pairs")
        if P[:, 0].min() < 0 or P[:, 0].max() >= integrals.shape[0]:
            raise ValueError("pair row outside the integrals")
        if P[:, 1:].min() < 0 or P[:, 3].max() > self.l_ej_max:
            raise ValueError("partial waves must be >= 0 and l_ej <= l_eject_max")
        k_i, k_f, k_ej = (np.broadcast_to(np.asarray(x, dtype=np.float64), (n,)) for x in (k_i, k_f, k_ej))
        ok = (k_i > 0) & (k_f > 0) & (k_ej > 0)
        scale = np.where(ok, (2.0 / math.pi) / np.where(ok, k_i * k_f, 1.0), 0.0)
        if self.norm:
            scale = scale * np.where(ok, math.sqrt(2.0 / math.pi) / np.where(ok, k_ej, 1.0), 0.0)
        types = _type_ids(self.types, P[:, 1:], self.device)
        lib = _hip.load_library()
        d_pairs = _pairs_device(P[:, 0], types, P[:, 2], scale, self.device)
        a = torch.empty((n, len(self.combos), 4), dtype=torch.float64, device=self.device)
        _hip.check(lib.dwba_amp_coeffs_f64(n, _hip.ptr(d_pairs), len(self.combos), self.L_max, _hip.ptr(integrals),
                                           _hip.ptr(self.types.d_coef), _hip.ptr(self.types.d_phase), _hip.ptr(a),
                                           _hip.stream_handle()), "dwba_amp_coeffs_f64")
        return a, d_pairs

    def coherent_sums(self, integrals: torch.Tensor, pairs, k_i, k_f, k_ej, slot, n_slots: int | None = None):
        """Coefficients summed over the pairs of each slot (e.g. slot = (l_ej,
        l_f) over l_i, the SDCS amplitudes of ionization.py:631-642), pairs of
        a slot in the order given.  Returns (n_slots, n_combos, 2) complex128."""
        slot = np.asarray(slot, dtype=np.int64).reshape(-1)
        P = np.asarray(pairs, dtype=np.int64).reshape(-1, 4)
        if slot.shape[0] != P.shape[0]:
            raise ValueError("one slot per pair")
        ns = int(slot.max()) + 1 if n_slots is None else int(n_slots)
        if slot.min() < 0 or slot.max() >= ns or ns > 65535:
            raise ValueError("slot outside [0, n_slots), n_slots <= 65535")
        if len(self.combos) > 64:
            raise ValueError("coherent sums take at most 64 projection pairs")
        n = P.shape[0]
        order = np.argsort(slot, kind="stable")
        k_i, k_f, k_ej = (np.broadcast_to(np.asarray(x, dtype=np.float64), (n,))[order] for x in (k_i, k_f, k_ej))
        a, d_pairs = self._coeffs(integrals, P[order], k_i, k_f, k_ej)
        first = np.zeros(ns + 1, dtype=np.int32)
        np.add.at(first, slot[order] + 1, 1)
        d_first = torch.as_tensor(np.cumsum(first).astype(np.int32), device=self.device)
        nc = len(self.combos)
        zeros = (ctypes.c_int32 * nc)()
        F = torch.empty((ns, nc, 1, 2), dtype=torch.float64, device=self.device)
        G = torch.empty_like(F)
        lib = _hip.load_library()
        _hip.check(lib.dwba_amp_sum_f64(ns, _hip.ptr(d_first), _hip.ptr(d_pairs), nc, zeros, zeros, 0, 0, 1, None,
                                        _hip.ptr(a), _hip.ptr(F), _hip.ptr(G), _hip.stream_handle()),
                   "dwba_amp_sum_f64")
        return torch.view_as_complex(torch.stack([F[:, :, 0], G[:, :, 0]], dim=2))


def _dict_row(I_dir, I_ex):
    keys = [int(k) for k in list(I_dir) + list(I_ex)]
    if not keys or min(keys) < 0 or max(keys) > 63:
        raise ValueError("I_L dicts need multipoles in [0, 63]")
    L_max = max(keys)
    row = np.zeros((1, 2, L_max + 1))
    for L, v in I_dir.items():
        row[0, 0, int(L)] = float(v)
    for L, v in I_ex.items():
        row[0, 1, int(L)] = float(v)
    return row, L_max


def calculate_amplitude_contribution(theta_grid, I_L_dir, I_L_exc, l_i, l_f, ki, kf, L_target_i, L_target_f,
                                     M_target_i, M_target_f) -> Amplitudes:
    """f(theta), g(theta) of one channel pair (l_i -> l_f) and projection pair
    (M_i, M_f), computed on the GPU (dwba_coupling.py:285-504)."""
    row, L_max = _dict_row(I_L_dir, I_L_exc)
    asm = AmplitudeAssembler(theta_grid, L_target_i, L_target_f, L_max, combos=[(M_target_i, M_target_f)])
    res = asm.excitation(torch.as_tensor(row, device=asm.device), [(0, l_i, l_f, 0)], [ki], [kf],
                         cross_sections=False)
    f = res.F[0, 0].cpu().numpy()
    g = res.G[0, 0].cpu().numpy()
    return Amplitudes(f, g)


def calculate_ionization_coefficients(I_L_dir, I_L_exc, l_i, l_f, l_eject, ki, kf, k_eject, L_target_i,
                                      M_target_i, M_target_f, include_eject_norm: bool = True):
    """Angle-free (e,2e) coefficients of one channel pair, on the GPU
    (dwba_coupling.py:507-634)."""
    if ki <= 0.0 or kf <= 0.0 or k_eject <= 0.0:
        return IonizationAmplitudeCoeffs(0j, 0j)
    if abs(M_target_f) > l_eject or abs(M_target_i) > L_target_i:
        return IonizationAmplitudeCoeffs(0j, 0j)          # no such state: every Clebsch-Gordan factor vanishes
    row, L_max = _dict_row(I_L_dir, I_L_exc)
    asm = IonizationAssembler(L_target_i, l_eject, L_max, include_eject_norm)
    c = asm.coefficients(torch.as_tensor(row, device=asm.device), [(0, l_i, l_f, l_eject)], [ki], [kf], [k_eject])
    k = asm.combos.index((int(M_target_i), int(M_target_f)))
    v = c[0, k].cpu().numpy()
    return IonizationAmplitudeCoeffs(complex(v[0]), complex(v[1]))


def dcs_dwba(theta_grid, f_theta, g_theta, ki, kf, L_i, N_equiv=1.0) -> np.ndarray:
    """(k_f/k_i) N_eq/(2L_i+1) [|f+g|^2/4 + 3|f-g|^2/4] (2 pi)^4, clipped at 0,
    on the GPU (sigma_total.py:147-177)."""
    d, _ = _dcs_tcs(theta_grid, f_theta, g_theta, ki, kf, L_i, N_equiv)
    return d


def integrate_dcs_over_angles(theta_grid, dcs_theta) -> float:
    """2 pi trapezoid(sin(theta) dcs, theta) on the GPU (sigma_total.py:179-217)."""
    dev = torch.device("cuda", _hip.ensure_init())
    th = torch.as_tensor(np.ascontiguousarray(theta_grid, dtype=np.float64), device=dev)
    d = torch.as_tensor(np.ascontiguousarray(dcs_theta, dtype=np.float64), device=dev)
    if d.shape != th.shape or th.dim() != 1:
        raise ValueError("theta_grid and dcs_theta must be 1-d of one length")
    tcs = torch.empty((1,), dtype=torch.float64, device=dev)
    _hip.check(_hip.load_library().dwba_sphere_trapz_f64(1, th.numel(), _hip.ptr(th), _hip.ptr(d), _hip.ptr(tcs),
                                                         _hip.stream_handle()), "dwba_sphere_trapz_f64")
    return float(tcs[0].cpu())


def _dcs_tcs(theta, f, g, ki, kf, Li, N_equiv):
    dev = torch.device("cuda", _hip.ensure_init())
    th = torch.as_tensor(np.ascontiguousarray(theta, dtype=np.float64), device=dev)
    F = torch.view_as_real(torch.as_tensor(np.ascontiguousarray(f, dtype=np.complex128), device=dev)).contiguous()
    G = torch.view_as_real(torch.as_tensor(np.ascontiguousarray(g, dtype=np.complex128), device=dev)).contiguous()
    nt = th.numel()
    if F.shape[0] != nt or G.shape[0] != nt:
        raise ValueError("f, g and theta differ in length")
    pref = (kf / ki) * (float(N_equiv) / float(2 * Li + 1)) if (ki > 0.0 and kf > 0.0) else 0.0
    d_pref = torch.tensor([pref], dtype=torch.float64, device=dev)
    dcs = torch.empty((1, nt), dtype=torch.float64, device=dev)
    tcs = torch.empty((1,), dtype=torch.float64, device=dev)
    lib = _hip.load_library()
    _hip.check(lib.dwba_dcs_f64(1, 1, nt, _hip.ptr(th), _hip.ptr(F), _hip.ptr(G), _hip.ptr(d_pref), _hip.ptr(dcs),
                                _hip.ptr(tcs), _hip.stream_handle()), "dwba_dcs_f64")
    return dcs[0].cpu().numpy(), float(tcs[0].cpu())
