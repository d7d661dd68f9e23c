"""CPU checks of the result records and their host arithmetic (dwba_amd.results)
against the reference's own stored excitation runs
(tests/golden/stored_results.json, taken from its results/*.json).

* Top-up: from an entry's per-l_i cross sections and its base sigma (stored
  sigma minus its stored top-up), excitation_topup reproduces the stored
  top-up key, value and final sigma.  Two stored runs predate the present
  top-up rules and are excluded by name: results_H_exc.json H_1s-2p (a
  Coulomb-Bethe tail applied without the monotonic-decay condition) and the
  119.093 eV entry of He+_1s-2s (a tail above the 20% acceptance gate kept).
* Units and the DCS integral: sigma_cm2 = sigma_au a0^2 for every entry;
  2 pi trapz(sin theta dcs_au_raw) is the stored sigma without its top-up
  (or with it, for runs written when the top-up rescaled the DCS).
* Format: an entry built here has the stored entry's keys in its order;
  save_results merges by key and writes indent=2 JSON that load_results reads back.
"""
import json
import math
import os

import numpy as np
import pytest

from dwba_amd import results as R

GOLD = os.path.join(os.path.dirname(__file__), "golden")
TARGETS = {"H_1s-2p": (0, 1), "H_1s-2s": (0, 0), "He+_1s-2p": (0, 1), "He+_1s-2s": (0, 0)}
EARLIER_RULES = {("results_H_exc.json:H_1s-2p", None), ("results_He+2s_exc.json:He+_1s-2s", 119.093)}


@pytest.fixture(scope="module")
def stored():
    with open(os.path.join(GOLD, "stored_results.json"), encoding="utf-8") as f:
        return json.load(f)


def _split(pw):
    return ({k: v for k, v in pw.items() if k[1:].isdigit()}, {k: v for k, v in pw.items() if not k[1:].isdigit()})


def test_topup_reproduces_stored_runs(stored):
    checked = 0
    for run, rows in stored["runs"].items():
        Li, Lf = TARGETS[run.split(":")[1]]
        if (run, None) in EARLIER_RULES:
            continue
        for e in rows:
            if (run, e["energy_eV"]) in EARLIER_RULES:
                continue
            pwL, top = _split(e["partial_waves"])
            sig, pw, meta = R.excitation_topup(pwL, e["sigma_cm2"] - sum(top.values()), Li, Lf)
            _, mine = _split(pw)
            assert set(mine) == set(top), (run, e["energy_eV"])
            for k in top:
                assert abs(mine[k] - top[k]) <= 1e-12 * abs(top[k])
            assert abs(sig - e["sigma_cm2"]) <= 1e-12 * e["sigma_cm2"]
            assert meta["topup_applied"] == bool(top)
            checked += 1
    assert checked >= 85


def test_units_and_dcs_integral(stored):
    n = 0
    for run, rows in stored["runs"].items():
        for e in rows:
            assert abs(e["sigma_cm2"] - R.sigma_au_to_cm2(e["sigma_au"])) <= 1e-12 * e["sigma_cm2"]
            if "dcs_au_raw" not in e:
                continue
            th = np.radians(np.asarray(e["theta_deg"]))
            y = np.sin(th) * np.asarray(e["dcs_au_raw"])
            base = R.sigma_au_to_cm2(2 * math.pi * float(np.sum(0.5 * (y[1:] + y[:-1]) * np.diff(th))))
            _, top = _split(e["partial_waves"])
            t = sum(top.values())
            assert (abs(base + t - e["sigma_cm2"]) <= 1e-10 * e["sigma_cm2"]
                    or abs(base - e["sigma_cm2"]) <= 1e-10 * e["sigma_cm2"]), (run, e["energy_eV"])
            n += 1
    assert n >= 15


def test_entry_keys_match_stored_format(stored):
    keys = stored["entry_keys"]["results_H2s_exc.json:H_1s-2s"]
    e = R.excitation_entry(10.0, 1.0, R.sigma_au_to_cm2(1.0), [0.0, 90.0], [1.0, 2.0], {"L0": 1.0},
                           runtime_meta={"L_max_integrals_used": 11, "L_max_projectile_used": 12,
                                         "L_dynamic_required": 12, "L_max_projectile_summed": 12,
                                         "n_projectile_partial_waves_summed": 13},
                           sigma_mtong_cm2=None, calibration_alpha=None, calibration_factor=1.0)
    assert list(e.keys()) == keys
    i = R.ionization_entry(50.0, 1.0, 2.0, 13.6, [(1.0, 2.0)], None, {"L0": 1.0})
    assert list(i.keys()) == ["energy_eV", "sigma_au", "sigma_cm2", "IP_eV", "sdcs", "tdcs", "partial_waves", "meta"]


def test_save_and_load_merge(tmp_path):
    p = R.get_json_path("H2p", "exc", root=tmp_path)
    assert p.name == "results_H2p_exc.json" and p.parent.name == "results"
    R.save_results(p.name, {"A": [1]}, root=tmp_path)
    R.save_results(p.name, {"B": [2]}, root=tmp_path)
    R.save_results(p.name, {"A": [3]}, root=tmp_path)
    assert R.load_results(p.name, root=tmp_path) == {"A": [3], "B": [2]}
    assert p.read_text().startswith('{\n  "A"')
    assert R.load_results("missing.json", root=tmp_path) == {}


def test_sdcs_quadrature_is_exact_for_polynomials():
    E = 37.0
    x, w, m = R.sdcs_energy_quadrature(E, 9, "gauss_legendre")
    assert m == "gauss_legendre" and x.size == 10
    half = E / 2
    for k in range(0, 19):
        exact = half ** (k + 1) / (k + 1) / R.HARTREE_TO_EV
        assert abs(float(np.dot(w, x ** k)) - exact) <= 1e-12 * exact
    x, w, m = R.sdcs_energy_quadrature(E, 9, "trapz_linear")
    assert m == "trapz_linear" and x[0] == 0.0 and x[-1] == half
    assert abs(float(np.dot(w, 3 * x + 1)) - (1.5 * half ** 2 + half) / R.HARTREE_TO_EV) <= 1e-12 * half ** 2
    assert R.sdcs_energy_quadrature(0.0, 9)[2] == "none"


def test_sdcs_tail_cases():
    assert R.sdcs_high_l_tail({"L0": 1.0, "L1": 0.5, "L2": 0.25}, 2.0)[1] == "none"       # fewer than 4
    assert R.sdcs_high_l_tail({f"L{l}": 1.0 + l for l in range(6)}, 10.0)[1] == "none"    # not decaying
    q = 0.5
    pl = {f"L{l}": q ** l for l in range(8)}
    t, method, d = R.sdcs_high_l_tail(pl, 2.0)
    geo = q ** 7 * q / (1 - q)
    assert method in ("geometric", "power_law") and 0 < t <= geo * (1 + 1e-12)
    assert abs(d["geometric_q"] - q) < 1e-15
    big = R.sdcs_high_l_tail(pl, 1e-6)
    assert big[1] == "rejected" and big[0] == 0.0


def test_ionization_total_weights():
    x, w, _ = R.sdcs_energy_quadrature(20.0, 5)
    s = np.full(x.size, 2.0)
    sig, cm2, table, pw = R.ionization_total(x, w, list(s), [{"L0": 1.0, "L1": 2.0}] * x.size)
    assert abs(sig - 2.0 * w.sum()) <= 1e-15 * sig
    assert abs(cm2 - R.sigma_au_to_cm2(sig)) <= 1e-15 * cm2
    assert abs(pw["L1"] - R.sigma_au_to_cm2(2.0 * w.sum())) <= 1e-15 * pw["L1"]
    assert len(table) == x.size and abs(table[0][1] - R.sigma_au_to_cm2(2.0 / R.HARTREE_TO_EV)) < 1e-30
