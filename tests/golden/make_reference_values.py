"""Writes ``reference_values.json``: data the reference's own tests, notebooks and
benchmark CSVs hold, read as text (nothing of the reference is imported).

Sources (paths relative to the reference checkout):
  test/test_energy_alignment.py:30-58   Julia golden energies (ATOL 1e-9, :27)
  test/test_core.py:173                 Haldane 2x3 E0
  benchmark/benchmark_data/benchmark_raw_latest_py.csv   sizes, orbit counts, sector dims, E0
  doc/design.ipynb cells 5, 12, 14, 16, 22, 24, 27       printed known-answer values
Run once where the reference is mounted: ``python tests/golden/make_reference_values.py``.
"""

from __future__ import annotations

import ast
import csv
import json
import re
import sys
from pathlib import Path

REF = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
OUT = Path(__file__).resolve().parent / "reference_values.json"


def energy_tables() -> dict:
    src = (REF / "test/test_energy_alignment.py").read_text()
    tree = ast.parse(src)
    out = {}
    for node in tree.body:
        if isinstance(node, ast.Assign) and isinstance(node.targets[0], ast.Name):
            name = node.targets[0].id
            if name.startswith("JULIA_REF_") or name == "ATOL":
                val = ast.literal_eval(node.value)
                if isinstance(val, dict):
                    val = {f"{k[0]},{k[1]}": v for k, v in val.items()}
                out[name] = val
    return out


def csv_rows() -> list[dict]:
    rows = []
    with open(REF / "benchmark/benchmark_data/benchmark_raw_latest_py.csv") as f:
        for r in csv.DictReader(f):
            rows.append({k: r[k] for k in r})
    return rows


def test_core_e0() -> float:
    src = (REF / "test/test_core.py").read_text()
    m = re.search(r"assertAlmostEqual\(vals\[0\], (-?[0-9.]+), places=9\)", src)
    return float(m.group(1))


def notebook_outputs() -> dict:
    nb = json.loads((REF / "doc/design.ipynb").read_text())
    text = {}
    for i, c in enumerate(nb["cells"]):
        if c["cell_type"] != "code":
            continue
        out = []
        for o in c.get("outputs", []):
            if "text" in o:
                out.append("".join(o["text"]))
        text[i] = "".join(out)
    kat = {}
    g = re.findall(r"'0b([01]+)'", text[5])
    kat["gosper_3_of_6"] = [int(b, 2) for b in g]
    kat["colex_rank_4_2"] = [[int(a), int(b)] for a, b in re.findall(r"^\s+(\d+) = \d+\s+bits=.*rank = (\d+)$",
                                                                       text[12], re.M)]
    m = re.search(r"shapes: \((\d+),\) \((\d+),\)\s+n_right = (\d+)", text[14])
    kat["lin_table_12_3"] = {"left": int(m.group(1)), "right": int(m.group(2)), "n_right": int(m.group(3)),
                             "mismatches": int(re.search(r"mismatches over all 220 masks: (\d+)", text[14]).group(1))}
    m16 = re.search(r"n_orbits = (\d+) \| full basis C\(12,3\) = (\d+)", text[16])
    unset = int(re.search(r"unset\(-1\) entries = (\d+)", text[16]).group(1))
    canon = re.search(r"get_canonical\(rep\[3\] = (\d+) \) = \((\d+), (\d+), \(([-0-9.+j]+)\)\)", text[16])
    kat["haldane_2x3_catalog"] = {"n_orbits": int(m16.group(1)), "n_total": int(m16.group(2)), "unset": unset,
                                  "rep3": int(canon.group(1)), "canonical": [int(canon.group(2)), int(canon.group(3))]}
    dim = re.search(r"sector dim = (\d+) \| nnz = (\d+)", text[22])
    e = re.search(r"E0\(k=\(0,0\)\) = (-?[0-9.]+)", text[22])
    dense = re.search(r"dense\s+: \[([-0-9. ]+)\]", text[22])
    kat["haldane_2x3_k0_block"] = {"dim": int(dim.group(1)), "nnz": int(dim.group(2)), "E0": float(e.group(1)),
                                   "lowest3_8dp": [float(x) for x in dense.group(1).split()]}
    t = re.search(r"projection table entries = (\d+)", text[24])
    kat["haldane_2x3_k0_table_entries"] = int(t.group(1))
    ph = re.search(r"θ_x = 0.25·2π: label \(1, 0\), perm_phases = \[([^\]]+)\]", text[27], re.S)
    vals = [complex(v.replace(" ", "")) for v in re.findall(r"[-0-9.]+[+-][0-9.]+j", ph.group(1))]
    kat["haldane_2x2_flux_phases_op2_theta_0.25"] = [[v.real, v.imag] for v in vals]
    return kat


def main() -> None:
    data = {
        "energy_alignment": energy_tables(),
        "test_core_haldane_2x3_E0": test_core_e0(),
        "benchmark_csv_latest": csv_rows(),
        "design_notebook": notebook_outputs(),
    }
    OUT.write_text(json.dumps(data, indent=1, sort_keys=True))
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
