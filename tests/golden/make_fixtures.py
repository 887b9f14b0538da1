"""Extract the reference's own stored results into a small JSON fixture.

Reads (as data, in this container) the benchmark CSVs the reference keeps under
benchmark/benchmark_data/ and the printed outputs of doc/design.ipynb and
doc/charge_pump.ipynb, and writes tests/golden/reference_values.json.  Nothing of
the reference is imported or run.
"""
import csv
import json
import os
import re

REF = "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_values.json")


def csv_rows(name):
    with open(os.path.join(REF, "benchmark", "benchmark_data", name)) as f:
        return list(csv.DictReader(f))


def main():
    data = {"benchmark_raw": [], "benchmark_raw_py_scan": [], "benchmark_raw_single": []}
    for key in data:
        for r in csv_rows(key + ".csv"):
            data[key].append({
                "model": r["model"], "label": r["label"], "n_site": int(r["n_site"]),
                "n_filled": int(r["n_filled"]), "full_dim": int(r["full_dim"]),
                "n_orbits": int(r["n_orbits"]), "n_group": int(r["n_group"]),
                "sector_dim": int(r["sector_dim"]), "mode": r["mode"], "energy": float(r["energy"]),
                "elapsed_s": float(r["elapsed_s"]),
            })
    nb = json.load(open(os.path.join(REF, "doc", "design.ipynb")))
    outs = {}
    for i, c in enumerate(nb["cells"]):
        txt = ""
        for o in c.get("outputs", []):
            t = o.get("text") or o.get("data", {}).get("text/plain") or ""
            txt += "".join(t)
        outs[i] = txt
    cell22 = outs[22]
    ranks = [int(x) for x in re.findall(r"rank = (\d+)", cell22)]
    masks = [int(x) for x in re.findall(r"^\s+(\d+) = ", cell22, re.M)]
    cell26 = outs[26]
    cell43 = outs[43]
    data["design_notebook"] = {
        "colex_n4_k2": {"masks": masks, "ranks": ranks},
        "lin_index_mismatches_12_3": int(re.search(r"mismatches over 220 masks: (\d+)", outs[24]).group(1)),
        "haldane_2x3": {
            "n_orbits": int(re.search(r"n_orbits = (\d+)", cell26).group(1)),
            "full_basis": int(re.search(r"C\(12,3\) = (\d+)", cell26).group(1)),
            "unset": int(re.search(r"unset = (\d+)", cell26).group(1)),
            "rep4_mask": int(re.search(r"get_canonical\(rep\[4\] = (\d+)\)", cell26).group(1)),
            "rep4_canonical": [int(x) for x in re.search(r"\(repr = (\d+), g_idx = (\d+)", cell26).groups()],
            "sector_dim": int(re.search(r"sector dim = (\d+)", cell43).group(1)),
            "nnz_H": int(re.search(r"nnz\(H\) = (\d+)", cell43).group(1)),
            "table_entries": int(re.search(r"projection table entries = (\d+)", cell43).group(1)),
        },
    }
    # doc/charge_pump.ipynb: bosonic Haldane 2x3, 3 bosons, 17 fluxes 0:1/16:1 along x
    # (cells 4, 5: flow and pump, polarization along y) and a 5x5 flux torus (cell 7)
    nb = json.load(open(os.path.join(REF, "doc", "charge_pump.ipynb")))
    pouts = {}
    for i, c in enumerate(nb["cells"]):
        pouts[i] = "".join("".join(o.get("text") or o.get("data", {}).get("text/plain") or "")
                           for o in c.get("outputs", []))
    num = r"(-?[0-9.]+(?:e-?[0-9]+)?)"
    flow = re.search(r"Scanned (\d+) sectors at (\d+) flux points", pouts[4])
    charges = re.search(r"Branch charges: \[" + num + ", " + num + r"\]", pouts[5])
    final = re.findall(r"sector = \((\d+), (\d+)\), level = (\d+)", pouts[5])
    data["charge_pump_notebook"] = {
        "model": "haldane_boson_FCI 2x3 (test/bosonic_fci.jl params), filling 1/4",
        "flow": {"n_sectors": int(flow.group(1)), "n_flux": int(flow.group(2))},
        "pump": {
            "pumped_charges": [float(charges.group(1)), float(charges.group(2))],
            "min_gap": float(re.search(r"Minimum gap: " + num, pouts[5]).group(1)),
            "selected_states_final": [[int(a), int(b), int(lv)] for a, b, lv in final],
        },
        "chern": {
            "chern_number": float(re.search(r"Manifold Chern number: " + num, pouts[7]).group(1)),
            "min_gap": float(re.search(r"Minimum gap: " + num, pouts[7]).group(1)),
            "min_link_det": float(re.search(r"minimum link determinant: " + num, pouts[7]).group(1)),
        },
    }
    json.dump(data, open(OUT, "w"), indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
