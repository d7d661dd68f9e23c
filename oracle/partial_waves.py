"""Oracle: the projectile partial-wave loops around the integrals.

Test infrastructure only.  Restates, step by step, on integrals given as a
lookup (the integral call itself is checked elsewhere):

  excitation_loop   the GPU branch of compute_total_excitation_cs,
                    driver.py:877-1152 (l_i loop, l_f parity range, per-(M_i, M_f)
                    amplitude sums, numerical-safety checks, per-angle DCS
                    stability stop), and the cross section of :1232-1260
  ionization_loop   the partial-wave loop of _compute_sdcs_at_energy,
                    ionization.py:526-705 (incoherent per-l_i sums, coherent
                    sums per (l_ej, l_f, M_i, M_f) updated incrementally, the
                    3-in-a-row relative-change stop), before the high-L top-up

with the amplitude routines of oracle.amplitude.
"""
from __future__ import annotations

import math
from collections import deque

import numpy as np

from . import amplitude as OA

A0_SQUARED_CM2 = 2.8002852e-17          # sigma_total.py:66


def excitation_loop(theta, integrals, has_chi_i, has_chi_f, k_i, k_f, Li, Lf, L_int, L_max_proj, N_equiv=1.0):
    """integrals(l_i, l_f) -> (I_dir dict, I_ex dict); has_chi_i(l), has_chi_f(l)
    -> bool.  Returns dict(F, G (per (Mi, Mf) in loop order), dcs, sigma_au,
    partial_waves (cm^2), stop_reason, stop_L, skipped)."""
    combos = [(mi, mf) for mi in range(-Li, Li + 1) for mf in range(-Lf, Lf + 1)]
    z = lambda: np.zeros_like(theta, dtype=complex)
    total = {c: [z(), z()] for c in combos}
    pw, history = {}, deque(maxlen=4)
    stop_reason = stop_L = None
    skipped = 0
    par = (Li + Lf) % 2
    for l_i in range(L_max_proj + 1):
        if not has_chi_i(l_i):
            skipped += 1
            continue
        li_amp = {c: [z(), z()] for c in combos}
        any_valid = False
        for l_f in range(0, max(L_max_proj, l_i + L_int) + 1):
            if (l_i + l_f) % 2 != par or not has_chi_f(l_f):
                continue
            any_valid = True
            d, e = integrals(l_i, l_f)
            for (mi, mf) in combos:
                f, g = OA.amplitude_contribution(theta, d, e, l_i, l_f, k_i, k_f, Li, Lf, mi, mf)
                li_amp[(mi, mf)][0] += f
                li_amp[(mi, mf)][1] += g
        if any_valid:
            dl = sum(OA.dcs_dwba(theta, fg[0], fg[1], k_i, k_f, Li, N_equiv) for fg in li_amp.values())
            pw[f"L{l_i}"] = OA.integrate_dcs(theta, dl) * A0_SQUARED_CM2
        add = True
        for fg in li_amp.values():
            if not (np.all(np.isfinite(fg[0])) and np.all(np.isfinite(fg[1]))):
                stop_reason, add = "Numerical failure: NaN/Inf in l_i amplitudes", False
                break
        if stop_reason is None:
            for fg in total.values():
                if not (np.all(np.isfinite(fg[0])) and np.all(np.isfinite(fg[1]))):
                    stop_reason, add = "Numerical failure: NaN/Inf in accumulated amplitudes", False
                    break
        if add:
            for c in combos:
                total[c][0] += li_amp[c][0]
                total[c][1] += li_amp[c][1]
        if stop_reason:
            stop_L = l_i
            break
        snap = sum(OA.dcs_dwba(theta, fg[0], fg[1], k_i, k_f, Li, N_equiv) for fg in total.values())
        if l_i >= 5:
            history.append((l_i, snap.copy()))
            if len(history) >= 4:
                old = history[-4][1]
                with np.errstate(divide="ignore", invalid="ignore"):
                    ch = np.abs(snap - old) / (snap + 1e-50)
                    ch[~np.isfinite(ch)] = 0.0
                if np.max(ch) < 0.01 and l_i > 15:
                    stop_reason, stop_L = "Converged", l_i
                    break
    dcs = sum(OA.dcs_dwba(theta, fg[0], fg[1], k_i, k_f, Li, N_equiv) for fg in total.values())
    return dict(combos=combos, F=np.array([total[c][0] for c in combos]), G=np.array([total[c][1] for c in combos]),
                dcs=dcs, sigma_au=OA.integrate_dcs(theta, dcs), partial_waves=pw, stop_reason=stop_reason,
                stop_L=stop_L, skipped=skipped)


def ionization_loop(integrals, has_chi_i, has_chi_ej, has_chi_f, k_i, k_scatt, k_eject, l_i_target, l_eject_max,
                    L_max, L_max_proj, l_f_max_scatt, N_equiv=1.0, convergence_threshold=0.01):
    """integrals(l_i, l_ej, l_f) -> (I_dir dict, I_ex dict).  Returns
    dict(sigma_au, partial_L (a.u.), coherent {(l_ej, l_f, Mi, Mf): [f, g]}, last_l_i)."""
    Li = l_i_target
    pref = (k_scatt * k_eject / k_i) * (float(N_equiv) / float(2 * Li + 1)) * (2 * math.pi) ** 4

    def spin(f, g):
        return 0.25 * abs(f + g) ** 2 + 0.75 * abs(f - g) ** 2

    amp, partial = {}, {}
    total, prev, small, last = 0.0, 0.0, 0, -1
    for l_i in range(L_max_proj + 1):
        if not has_chi_i(l_i):
            break
        sigma_L = 0.0
        for l_ej in range(l_eject_max + 1):
            if not has_chi_ej(l_ej):
                continue
            lf_max = min(max(L_max_proj, l_i + L_max), l_f_max_scatt)
            par = (Li + l_ej) % 2
            combo = 0.0
            for l_f in range(0, lf_max + 1):
                if (l_i + l_f) % 2 != par or not has_chi_f(l_f):
                    continue
                d, e = integrals(l_i, l_ej, l_f)
                for mi in range(-Li, Li + 1):
                    for mf in range(-l_ej, l_ej + 1):
                        f, g = OA.ionization_coefficients(d, e, l_i, l_f, l_ej, k_i, k_scatt, k_eject, Li, mi, mf)
                        if f == 0.0 and g == 0.0:
                            continue
                        combo += spin(f, g)
                        key = (l_ej, l_f, mi, mf)
                        if key not in amp:
                            amp[key] = [f, g]
                            total += spin(f, g)
                        else:
                            of, og = amp[key]
                            old = spin(of, og)
                            amp[key] = [of + f, og + g]
                            total += spin(of + f, og + g) - old
            sigma_L += pref * combo
        partial[f"L{l_i}"] = sigma_L
        last = l_i
        sig = pref * total
        if l_i > 3 and prev > 1e-40:
            if abs(sig - prev) / prev < convergence_threshold:
                small += 1
                if small >= 3:
                    break
            else:
                small = 0
        prev = sig
    return dict(sigma_au=pref * total, partial_L=partial, coherent=amp, last_l_i=last)
