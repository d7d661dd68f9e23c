"""Write tests/golden/stored_results.json: the fields of the reference's
stored excitation runs (results/results_*_exc.json, data files the reference
ships) that the result-format tests use -- per entry the energy, sigma (a.u.
and cm^2), partial waves and runtime metadata, and for the first three
entries of each run also theta_deg and dcs_au_raw.  The key order of one full
entry of each file is kept under "entry_keys".

Usage:  python tests/golden/make_stored_results.py [/root/reference]
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def main(ref_root):
    out = {"runs": {}, "entry_keys": {}}
    d = os.path.join(ref_root, "results")
    for fn in sorted(os.listdir(d)):
        if not (fn.startswith("results_") and fn.endswith("_exc.json")):
            continue
        with open(os.path.join(d, fn), encoding="utf-8") as f:
            data = json.load(f)
        for key, entries in data.items():
            out["entry_keys"][f"{fn}:{key}"] = list(entries[0].keys())
            rows = []
            for i, e in enumerate(entries):
                r = {k: e[k] for k in ("energy_eV", "sigma_au", "sigma_cm2", "partial_waves")}
                r["runtime"] = (e.get("meta") or {}).get("runtime") or {}
                if i < 3:
                    r["theta_deg"] = e["theta_deg"]
                    r["dcs_au_raw"] = e["dcs_au_raw"]
                rows.append(r)
            out["runs"][f"{fn}:{key}"] = rows
    with open(os.path.join(HERE, "stored_results.json"), "w", encoding="utf-8") as f:
        json.dump(out, f)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
