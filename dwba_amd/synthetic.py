"""Seeded synthetic inputs of the benchmark workloads (host side, numpy).

* grid: r_j = 1e-5 exp(j h), h = ln(1e8)/(N-1) (r_max = 1000 a0), Simpson
  weights;
* bound states: u_1s = 2 r e^{-r}, u_2p = r^2 e^{-r/2} / (2 sqrt 6),
  Slater u = (2 zeta)^{3/2}/sqrt(2) r e^{-zeta r};
* continuum waves: chi_l(r) = sin(k r - l pi/2 + delta) (1 - e^{-r})^{l+1};
* direct densities (chi_f chi_i, u_f u_i), exchange (u_f chi_i, chi_f u_i).

The objects returned quack like the reference's BoundOrbital / ContinuumWave
(bound_states.py:59-89, continuum.py:68-119).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .grid import RadialGrid, exp_grid


@dataclass
class Bound:
    """BoundOrbital stand-in: u_of_r (+ l, energy_au, n_index)."""
    u_of_r: np.ndarray
    l: int = 0
    energy_au: float = -0.5
    n_index: int = 1


@dataclass
class Wave:
    """ContinuumWave stand-in with unit asymptotic amplitude."""
    l: int
    k_au: float
    chi_of_r: np.ndarray
    phase_shift: float = 0.0
    r_match: float = 0.0
    idx_match: int = -1
    eta: float = 0.0
    sigma_l: float = 0.0

    @property
    def u_of_r(self) -> np.ndarray:
        return self.chi_of_r


def u1s(r):
    return 2.0 * r * np.exp(-r)


def u2p(r):
    return r * r * np.exp(-r / 2) / (2.0 * np.sqrt(6.0))


def slater(r, zeta=0.6):
    return (2 * zeta) ** 1.5 / np.sqrt(2.0) * r * np.exp(-zeta * r)


_WENV: dict = {}


def _wave_envelope(r, l):
    """(1 - e^{-r})^{l+1}, cached per (grid, l)."""
    key = (id(r), len(r), float(r[0]), float(r[-1]), int(l))
    hit = _WENV.get(key)
    if hit is not None and hit[0] is r:
        return hit[1]
    v = (1.0 - np.exp(-r)) ** (l + 1)
    if len(_WENV) > 8192:
        _WENV.clear()
    _WENV[key] = (r, v)
    return v


def wave_values(r, k, l, delta):
    return np.sin(k * r - l * np.pi / 2 + delta) * _wave_envelope(r, l)


_ENV_CACHE: dict = {}


def _envelope_index(r, l, tol):
    """First index from which 1 - (1-e^{-r})^{l+1} < tol (the deficit decreases with r)."""
    key = (id(r), len(r), float(r[0]), float(r[-1]), int(l), tol)
    hit = _ENV_CACHE.get(key)
    if hit is not None and hit[0] is r:
        return hit[1]
    deficit = -np.expm1((l + 1) * np.log1p(-np.exp(-r)))
    ok = np.nonzero(deficit < tol)[0]
    v = int(ok[0]) if len(ok) else len(r)
    if len(_ENV_CACHE) > 4096:
        _ENV_CACHE.clear()
    _ENV_CACHE[key] = (r, v)
    return v


def match_index(r, k, l, tol=1e-12) -> int:
    """Match point of a synthetic wave: first grid index beyond the turning
    point (l + 1/2)/k where the envelope deficit 1 - (1-e^{-r})^{l+1} < tol."""
    r_turn = (l + 0.5) / k
    i = max(_envelope_index(r, l, tol), int(np.searchsorted(r, r_turn, side="right")))
    return i if i < len(r) else len(r) - 1


def coulomb_eta_sigma(k, l, z_ion):
    """Sommerfeld parameter eta = -z_ion / k of a wave in the field of an ion
    of charge z_ion and its Coulomb phase sigma_l = arg Gamma(l + 1 + i eta)."""
    if not z_ion:
        return 0.0, 0.0
    from scipy.special import loggamma
    eta = -float(z_ion) / float(k)
    return eta, float(np.imag(loggamma(l + 1 + 1j * eta)))


def make_wave(r, k, l, delta, with_match=True, z_ion=0.0) -> Wave:
    """Synthetic wave chi_l; with z_ion != 0 it carries the Coulomb eta and
    sigma_l of an ionic target (its shape is the same synthetic form: only
    the phase metadata the quadrature reads changes)."""
    eta, sigma = coulomb_eta_sigma(k, l, z_ion)
    return Wave(l=int(l), k_au=float(k), chi_of_r=wave_values(r, k, l, delta), phase_shift=float(delta),
                idx_match=match_index(r, k, l) if with_match else -1, eta=eta, sigma_l=sigma)


@dataclass
class PairDensities:
    """alpha/beta-ready densities of one channel pair on the full grid."""
    rho1_dir: np.ndarray
    rho2_dir: np.ndarray
    rho1_ex: np.ndarray
    rho2_ex: np.ndarray


def plain_pair(grid: RadialGrid, seed: int = 0, k_range=(0.5, 8.0), l_max: int = 40):
    """Single channel pair of the kernel-only workload: H 1s -> 2p bound pair,
    waves with k ~ U[k_range], delta ~ U[0, pi], l ~ U{0..l_max}.  Returns the
    Simpson-weighted (alpha, beta) of direct and exchange and the draws."""
    rng = np.random.default_rng(seed)
    r, w = grid.r, grid.w_simpson
    ki, kf = rng.uniform(*k_range, size=2)
    li, lf = rng.integers(0, l_max + 1, size=2)
    di, df = rng.uniform(0.0, np.pi, size=2)
    chi_i = wave_values(r, ki, li, di)
    chi_f = wave_values(r, kf, lf, df)
    ui, uf = u1s(r), u2p(r)
    ad = w * chi_f * chi_i
    bd = w * uf * ui
    ae = w * uf * chi_i
    be = w * chi_f * ui
    draws = dict(k_i=ki, k_f=kf, l_i=int(li), l_f=int(lf), delta_i=di, delta_f=df)
    return ad, bd, ae, be, draws


def kernel_only_workload(n: int = 15000, seed: int = 0):
    """Kernel-only single pair: N = 15000 exp grid on [1e-5, 1000]."""
    g = exp_grid(n, 1e-5, 1e3)
    return g, plain_pair(g, seed)


def h_potentials(r):
    """V_core = -1/r (Z = 1) and the static distorting potential of H(1s),
    U = -(1 + 1/r) e^{-2r} (V_core + Hartree of the 1s density)."""
    V = -1.0 / r
    U = -(1.0 + 1.0 / r) * np.exp(-2.0 * r)
    return V, U


def excitation_scan(n_energies=64, n_pairs=400, n_range=(4000, 15000), l_max=40, k_range=(0.5, 8.0),
                    seed=0, with_match=True, z_ion=0.0):
    """Batched precision-sweep workload: per energy its own exp grid
    N ~ U{n_range} on [1e-5, 1000] a0, (k_i, k_f) ~ U[k_range] per energy,
    n_pairs channel pairs with (l_i, l_f) ~ U{0..l_max}, delta ~ U[0, pi]
    per wave; bound pair u_1s -> u_2p; waves carry match points.  z_ion != 0:
    the waves carry the Coulomb eta and sigma_l of an ion of that charge
    (Coulomb-stepped Filon phase nodes)."""
    from .batch import EnergyInput
    rng = np.random.default_rng(seed)
    out = []
    for e in range(n_energies):
        n = int(rng.integers(n_range[0], n_range[1] + 1))
        g = exp_grid(n, 1e-5, 1e3)
        r = g.r
        ki, kf = rng.uniform(*k_range, size=2)
        V, U = h_potentials(r)
        pairs = []
        for _ in range(n_pairs):
            li, lf = rng.integers(0, l_max + 1, size=2)
            di, df = rng.uniform(0.0, np.pi, size=2)
            pairs.append((make_wave(r, ki, li, di, with_match, z_ion), make_wave(r, kf, lf, df, with_match, z_ion)))
        out.append(EnergyInput(grid=g, V_core=V, U_i=U, bound_i=Bound(u1s(r), l=0, energy_au=-0.5),
                               bound_f=Bound(u2p(r), l=1, energy_au=-0.125), pairs=pairs, U_f=U))
    return out


def ionization_scan(n_points=48, n_triples=400, n_range=(6000, 15000), l_max=30, seed=0, k_i=2.0, z_ion=0.0):
    """Ionization-style workload: Slater bound state (zeta = 0.6) and three
    continuum waves (incident k_i, ejected k_e ~ U[0.2, 4], scattered k_f with
    k_f^2 = k_i^2 - k_e^2 - 2*0.18) per ejected-energy point on its own grid
    N ~ U{n_range}; n_triples (l_i, l_f, l_ej) ~ U{0..l_max} per point, drawn
    with l_i + l_f + l_ej even.  The ejected wave takes the place of u_f.
    z_ion != 0: Coulomb eta and sigma_l on every wave, as in excitation_scan."""
    from .batch import EnergyInput
    rng = np.random.default_rng(seed)
    out = []
    for e in range(n_points):
        n = int(rng.integers(n_range[0], n_range[1] + 1))
        g = exp_grid(n, 1e-5, 1e3)
        r = g.r
        ke = rng.uniform(0.2, 4.0)
        kf = np.sqrt(max(k_i ** 2 - ke ** 2 - 0.36, 0.04))
        V, U = h_potentials(r)
        pairs = []
        while len(pairs) < n_triples:
            li, lf, le = rng.integers(0, l_max + 1, size=3)
            if (li + lf + le) % 2:
                continue
            di, df, de = rng.uniform(0.0, np.pi, size=3)
            pairs.append((make_wave(r, k_i, li, di, z_ion=z_ion), make_wave(r, kf, lf, df, z_ion=z_ion),
                          make_wave(r, ke, le, de, z_ion=z_ion)))
        out.append(EnergyInput(grid=g, V_core=V, U_i=U, bound_i=Bound(slater(r, 0.6), l=0, energy_au=-0.18),
                               bound_f=pairs[0][2], pairs=pairs, U_f=U))
    return out


def backward_workload(n: int = 15000, n_pairs: int = 32, L_max: int = 20, seed: int = 0):
    """Gradient workload: N-point exp grid on [1e-5, 1000] a0 and n_pairs
    channel pairs as in the kernel-only workload (k ~ U[0.5, 8], delta ~
    U[0, pi], l ~ U{0..40}); even pairs are direct (rho_a = u_1s u_2p,
    rho_b = chi_f chi_i), odd pairs exchange (rho_a = u_2p chi_i,
    rho_b = chi_f u_1s).  Loss weights c_L ~ N(0, 1) seeded, shared by all
    pairs.  Returns (grid, rho_a (P, N), rho_b (P, N), g (P, L_max+1))."""
    g = exp_grid(n, 1e-5, 1e3)
    r = g.r
    rng = np.random.default_rng(seed)
    ui, uf = u1s(r), u2p(r)
    ra = np.empty((n_pairs, n))
    rb = np.empty((n_pairs, n))
    for p in range(n_pairs):
        ki, kf = rng.uniform(0.5, 8.0, size=2)
        li, lf = rng.integers(0, 41, size=2)
        di, df = rng.uniform(0.0, np.pi, size=2)
        ci, cf = wave_values(r, ki, li, di), wave_values(r, kf, lf, df)
        if p % 2 == 0:
            ra[p], rb[p] = ui * uf, ci * cf
        else:
            ra[p], rb[p] = uf * ci, cf * ui
    c = rng.standard_normal(L_max + 1)
    return g, ra, rb, np.tile(c, (n_pairs, 1))
