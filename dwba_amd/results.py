"""Result records in the reference's format, and the host arithmetic that
turns partial-wave sums into the numbers it stores.

  save_results / load_results / get_json_path   output_utils.py:27-222
      results/results_<run>_<exc|ion>.json, a dict {run key: [entries]}
      merged by key and written with indent=2
  excitation_entry   the per-energy entry of DW_main.py:2203-2223
  ionization_entry   the per-energy entry of DW_main.py:2446-2455
  excitation_topup   the high-L top-up of compute_total_excitation_cs
      (driver.py:1262-1420): Coulomb-Bethe for |L_f - L_i| = 1, a geometric
      (Born) tail otherwise, applied to the TCS only and rejected when it
      adds 20% or more
  sdcs_energy_quadrature, sdcs_high_l_tail, ionization_total
      the ejected-energy quadrature (ionization.py:215-263), the SDCS high-L
      tail (:266-352) and the SDCS integral with its per-l_i sums
      (:1048-1080)

These are a few scalar operations per energy on the partial-wave sums that
dwba_amd.partial_waves forms on the device; they run on the host, as in the
reference, and follow its formulas step by step so the stored numbers are
the same.
"""
from __future__ import annotations

import json
import math
from pathlib import Path
from typing import Dict, List, Optional, Tuple, Union

import numpy as np

A0_SQUARED_CM2 = 2.8002852e-17          # sigma_total.py:66
HARTREE_TO_EV = 27.211386245988         # grid.py:36
EV_TO_HARTREE = 1.0 / HARTREE_TO_EV
RESULTS_DIR = "results"


def sigma_au_to_cm2(x):
    return x * A0_SQUARED_CM2


def sigma_cm2_to_au(x):
    return x / A0_SQUARED_CM2


def ev_to_au(e):
    return np.asarray(e) * EV_TO_HARTREE if isinstance(e, np.ndarray) else e * EV_TO_HARTREE


# ------------------------------------------------------------------ files
def get_results_dir(root: Union[str, Path, None] = None) -> Path:
    """`results/` under `root` (default: the working directory), created if missing."""
    p = Path(root) / RESULTS_DIR if root is not None else Path(RESULTS_DIR)
    p.mkdir(parents=True, exist_ok=True)
    return p


def get_output_path(filename: Union[str, Path], root=None) -> Path:
    return get_results_dir(root) / Path(filename).name


def get_json_path(run_name: str, calc_type: str, root=None) -> Path:
    """results/results_<run_name>_<calc_type>.json (calc_type "exc" or "ion")."""
    return get_output_path(f"results_{run_name}_{calc_type}.json", root)


def load_results(filename: Union[str, Path], root=None) -> dict:
    """The stored dict, from results/ first, then the bare file name; {} when
    missing or unreadable."""
    base = Path(filename).name
    for p in (get_results_dir(root) / base, Path(base)):
        if p.exists():
            try:
                with open(p, "r", encoding="utf-8") as f:
                    return json.load(f)
            except Exception:
                return {}
    return {}


def save_results(filename: Union[str, Path], new_data_dict: dict, root=None) -> Path:
    """Merge `new_data_dict` into the stored dict (key by key) and write it
    back with indent=2.  Returns the path written."""
    base = Path(filename).name
    out = get_output_path(base, root)
    cur = load_results(base, root)
    cur.update(new_data_dict)
    with open(out, "w", encoding="utf-8") as f:
        json.dump(cur, f, indent=2)
    return out


# ------------------------------------------------------------------ excitation
def excitation_topup(partial_waves: Dict[str, float], sigma_cm2_base: float, L_target_i: int,
                     L_target_f: int) -> Tuple[float, Dict[str, float], dict]:
    """The top-up of driver.py:1262-1420 on the per-l_i cross sections (cm^2,
    keys "L{l}").  Returns (sigma_cm2, partial_waves with the top-up key when
    kept, {applied, type, value_cm2, tail_fraction})."""
    pw = dict(partial_waves)
    sig = float(sigma_cm2_base)
    applied, kind, value = False, None, 0.0
    ls = sorted(int(k[1:]) for k in pw if k.startswith("L") and not k.endswith("_excluded") and k[1:].isdigit())
    if abs(L_target_f - L_target_i) == 1:
        if len(ls) >= 3:
            L, Lm1, Lm2 = ls[-1], ls[-2], ls[-3]
            v, v1, v2 = pw.get(f"L{L}", 0), pw.get(f"L{Lm1}", 0), pw.get(f"L{Lm2}", 0)
            if v > 0 and v1 > v and v2 > v1 and L >= 3:
                q1, q2 = v / v1, v1 / v2
                if 0.5 < q1 < 0.98 and abs(q1 - q2) < 0.2:
                    tail = v * (np.log(2.0 * L + 1) / (L + 1)) * 2.0
                    if 1e-50 < tail < sigma_cm2_base * 0.5:
                        sig += tail
                        pw["coulomb_bethe_topup"] = tail
                        applied, kind, value = True, "Coulomb-Bethe", tail
    elif pw and len(ls) >= 3:
        for i in range(len(ls) - 1, 1, -1):
            v, v1, v2 = pw.get(f"L{ls[i]}", 0), pw.get(f"L{ls[i - 1]}", 0), pw.get(f"L{ls[i - 2]}", 0)
            if v > 0 and v1 > v and v2 > v1:
                q, qp = v / v1, v1 / v2
                if q < 0.95 and abs(q - qp) < 0.3:
                    tail = v * q / (1.0 - q)
                    if tail > 1e-50:
                        sig += tail
                        pw["born_topup"] = tail
                        applied, kind, value = True, "Born", tail
                        break
    frac = 0.0
    if applied and sigma_cm2_base > 0.0:
        frac = (sig - sigma_cm2_base) / sigma_cm2_base
        if frac >= 0.2:
            sig = float(sigma_cm2_base)
            pw.pop("coulomb_bethe_topup" if kind == "Coulomb-Bethe" else "born_topup", None)
            applied, kind, value, frac = False, None, 0.0, 0.0
    return sig, pw, {"topup_applied": applied, "topup_type": kind, "topup_value_cm2": float(value),
                     "topup_tail_fraction": float(frac)}


def excitation_runtime(result, L_max_integrals_used: int, L_max_projectile_used: int, topup: dict) -> dict:
    """The runtime metadata fields a partial-wave sum determines
    (driver.py:1422-1444)."""
    ls = sorted(int(k[1:]) for k in result.partial_waves if k.startswith("L") and k[1:].isdigit())
    return {
        "L_max_integrals_used": int(L_max_integrals_used),
        "L_max_projectile_used": int(L_max_projectile_used),
        "L_max_projectile_summed": int(ls[-1] if ls else -1),
        "n_projectile_partial_waves_summed": int(len(ls)),
        "n_projectile_partial_waves_skipped": int(result.skipped_li),
        "summation_stop_reason": result.stop_reason or "Reached configured/runtime L_max",
        "summation_stop_L": int(result.stop_L if result.stop_L is not None else L_max_projectile_used),
        **topup,
    }


def excitation_entry(energy_eV: float, sigma_au: float, sigma_cm2: float, theta_deg, dcs_au, partial_waves,
                     runtime_meta: Optional[dict] = None, run_meta: Optional[dict] = None,
                     sigma_mtong_cm2: Optional[float] = None, calibration_alpha: Optional[float] = None,
                     calibration_factor: float = 1.0) -> dict:
    """One energy of an excitation run, keys in the order DW_main.py:2203-2223
    writes them (the calibration fields are passed through: calibration is
    not part of this package)."""
    rt = dict(runtime_meta or {})
    dcs = None if dcs_au is None else [float(x) for x in np.asarray(dcs_au)]
    return {
        "energy_eV": float(energy_eV),
        "sigma_au": float(sigma_au),
        "sigma_cm2": float(sigma_cm2),
        "sigma_mtong_cm2": sigma_mtong_cm2,
        "calibration_alpha": calibration_alpha,
        "calibration_factor": calibration_factor,
        "theta_deg": None if theta_deg is None else [float(x) for x in np.asarray(theta_deg)],
        "dcs_au_raw": dcs,
        "dcs_au_calibrated": None if dcs is None else [x * calibration_factor for x in dcs],
        "partial_waves": {k: float(v) for k, v in partial_waves.items()},
        "L_max_integrals_used": rt.get("L_max_integrals_used"),
        "L_max_projectile_used": rt.get("L_max_projectile_used"),
        "L_dynamic_required": rt.get("L_dynamic_required"),
        "maxL_in_result": rt.get("L_max_projectile_summed"),
        "n_projectile_partial_waves_summed": rt.get("n_projectile_partial_waves_summed"),
        "meta": {**(run_meta or {}), "runtime": rt},
    }


# ------------------------------------------------------------------ ionization
def sdcs_energy_quadrature(E_total_final_eV: float, n_energy_steps: int, method: str = "gauss_legendre"):
    """Nodes (eV) and weights (hartree) over E_ej in [0, (E - IP)/2]
    (ionization.py:215-263): Gauss-Legendre with max(4, n + 1) nodes, or the
    trapezoidal rule on max(2, n + 1) equally spaced nodes."""
    E_half = max(0.5 * float(E_total_final_eV), 0.0)
    if E_half <= 1e-12:
        return np.array([], dtype=float), np.array([], dtype=float), "none"
    m = str(method).strip().lower()
    if m in ("auto", "", "gauss", "legendre"):
        m = "gauss_legendre"
    if m not in ("gauss_legendre", "trapz_linear"):
        m = "gauss_legendre"
    if m == "gauss_legendre":
        x, w = np.polynomial.legendre.leggauss(max(4, int(n_energy_steps) + 1))
        return (0.5 * E_half * (x + 1.0)).astype(float), (0.5 * E_half * w * ev_to_au(1.0)).astype(float), m
    n = max(2, int(n_energy_steps) + 1)
    nodes = np.linspace(0.0, E_half, n, dtype=float)
    w = np.zeros_like(nodes)
    dx = np.diff(nodes)
    w[0], w[-1] = 0.5 * dx[0], 0.5 * dx[-1]
    if n > 2:
        w[1:-1] = 0.5 * (dx[:-1] + dx[1:])
    return nodes, w * ev_to_au(1.0), "trapz_linear"


def sdcs_high_l_tail(partial_L: Dict[str, float], sigma_base_au: float) -> Tuple[float, str, dict]:
    """The SDCS high-L tail of ionization.py:266-352: a power law
    C (L + 1/2)^p fitted to the last (up to 6) per-l_i values, or a geometric
    tail, the smaller accepted, rejected at 25% of the base."""
    ls = sorted(int(k[1:]) for k in partial_L if k.startswith("L"))
    diag = {"applied": 0.0, "tail_au": 0.0, "tail_fraction": 0.0, "fit_r2": np.nan, "fit_slope": np.nan,
            "geometric_q": np.nan}
    if len(ls) < 4:
        return 0.0, "none", diag
    vals = np.array([float(partial_L.get(f"L{L}", 0.0)) for L in ls], dtype=float)
    if not np.all(np.isfinite(vals)):
        return 0.0, "none", diag
    nt = min(6, len(ls))
    Lt = np.array(ls[-nt:], dtype=float)
    vt = np.array(vals[-nt:], dtype=float)
    if np.any(vt <= 0.0):
        return 0.0, "none", diag
    if int(np.sum(np.diff(vt) < 0.0)) < max(2, nt - 2):
        return 0.0, "none", diag
    cands = []
    x, y = np.log(Lt + 0.5), np.log(vt)
    slope, icpt = np.linalg.lstsq(np.vstack([x, np.ones_like(x)]).T, y, rcond=None)[0]
    ss_res = float(np.sum((y - (slope * x + icpt)) ** 2))
    ss_tot = float(np.sum((y - np.mean(y)) ** 2)) + 1e-30
    r2 = 1.0 - ss_res / ss_tot
    diag["fit_r2"], diag["fit_slope"] = float(r2), float(slope)
    if slope < -1.05 and r2 > 0.90:
        L0 = float(Lt[-1] + 0.5)
        t = float(np.exp(icpt)) * (L0 ** (slope + 1.0)) / (-(slope + 1.0))
        if np.isfinite(t) and t > 0.0:
            cands.append(("power_law", float(t)))
    if len(vt) >= 3:
        ratios = vt[1:] / np.maximum(vt[:-1], 1e-300)
        q = float(np.median(ratios[-3:]))
        diag["geometric_q"] = q
        if 0.0 < q < 0.98 and float(np.max(np.abs(ratios[-3:] - q))) < 0.12:
            t = float(vt[-1] * q / (1.0 - q))
            if np.isfinite(t) and t > 0.0:
                cands.append(("geometric", t))
    if not cands:
        return 0.0, "none", diag
    method, tail = min(cands, key=lambda kv: kv[1])
    frac = tail / max(float(sigma_base_au), 1e-30)
    diag["tail_au"], diag["tail_fraction"] = float(tail), float(frac)
    if frac >= 0.25:
        return 0.0, "rejected", diag
    diag["applied"] = 1.0
    return float(tail), method, diag


def ionization_total(nodes_eV, weights_au, sdcs_au: List[float], partials_per_step: List[Dict[str, float]]):
    """sigma = sum_k w_k SDCS(E_k) (a.u. and cm^2), the SDCS table
    [(E_ej eV, cm^2/eV)] and the per-l_i sums integrated over E_ej (cm^2)
    (ionization.py:1048-1080)."""
    w = np.asarray(weights_au, dtype=float)
    sig_au = float(np.dot(w, np.asarray(sdcs_au, dtype=float)))
    pw = None
    if partials_per_step:
        pw = {}
        keys = set()
        for p in partials_per_step:
            keys.update(p.keys())
        for k in sorted(keys):
            pw[k] = sigma_au_to_cm2(float(np.dot(w, np.array([p.get(k, 0.0) for p in partials_per_step]))))
    table = [(float(e), float(sigma_au_to_cm2(s * ev_to_au(1.0)))) for e, s in zip(nodes_eV, sdcs_au)]
    return sig_au, float(sigma_au_to_cm2(sig_au)), table, pw


def ionization_entry(energy_eV: float, sigma_au: float, sigma_cm2: float, IP_eV: float, sdcs_data, tdcs_data,
                     partial_waves, meta: Optional[dict] = None) -> dict:
    """One energy of an ionization run, keys in the order of DW_main.py:2446-2455."""
    return {
        "energy_eV": float(energy_eV),
        "sigma_au": float(sigma_au),
        "sigma_cm2": float(sigma_cm2),
        "IP_eV": float(IP_eV),
        "sdcs": None if sdcs_data is None else [[float(a), float(b)] for a, b in sdcs_data],
        "tdcs": tdcs_data,
        "partial_waves": partial_waves,
        "meta": meta or {},
    }


# ------------------------------------------------------------------ from the partial-wave sums
def excitation_result_entry(energy_eV: float, res, theta_grid, L_target_i: int, L_target_f: int,
                            L_max_integrals_used: int, L_max_projectile_used: int, run_meta: Optional[dict] = None,
                            **calibration) -> dict:
    """The stored entry of one energy from a partial_waves.ExcitationResult:
    the top-up on its per-l_i sums, sigma in a.u. and cm^2, the raw DCS on
    theta in degrees, the runtime fields (driver.py:1262-1455)."""
    sig_cm2, pw, top = excitation_topup(res.partial_waves, res.sigma_cm2, L_target_i, L_target_f)
    sig_au = sigma_cm2_to_au(sig_cm2) if top["topup_applied"] else res.sigma_au
    rt = excitation_runtime(res, L_max_integrals_used, L_max_projectile_used, top)
    return excitation_entry(energy_eV, sig_au, sig_cm2, np.degrees(np.asarray(theta_grid, dtype=float)), res.dcs, pw,
                            runtime_meta=rt, run_meta=run_meta, **calibration)


def ionization_point_sdcs(res) -> Tuple[float, Dict[str, float], dict]:
    """SDCS of one ejected energy from a partial_waves.IonizationResult with
    the high-L tail added as _compute_sdcs_at_energy does (ionization.py:706-727):
    (SDCS a.u., per-l_i contributions (+ "born_topup"), diagnostics)."""
    sdcs = float(res.sigma_au)
    partial = dict(res.partial_L)
    tail, method, diag = sdcs_high_l_tail(partial, sdcs)
    if tail > 0.0:
        sdcs += tail
        partial["born_topup"] = tail
    diag["method_code"] = float({"none": 0.0, "power_law": 1.0, "geometric": 2.0, "rejected": 3.0}.get(method, -1.0))
    return sdcs, partial, diag
