"""Writes the parity fixtures of the grid-map path.

1. For every synthetic case in tests/cases.py, writes its seeded inputs
   (PDBQT + GPF) under tests/golden/cases/<name>/.
2. Runs the reference autogrid4, built from its own sources by
   oracle/Makefile into oracle/_ref/autogrid4, on every case in a scratch
   directory and records in tests/golden/golden.json:
     - the exit status and whether the log ends in "Successful Completion";
     - per map file, the SHA-256 of its value lines (everything after the 6
       header lines) and the number of values;
     - the .xyz text and the .fld text without its "Created by" line;
     - the log lines consumers parse (cases.glg_record): the
       'Calculating ...' line, the per-plane table's header and each
       plane's number, z and percent, the extrema table with its values as
       printed, the clamp note and the completion line.

3. When oracle/_ref/autogrid4_bhtree is built (make -C oracle ref-bhtree),
   runs it on the tree-build cases (cases.BH_CASES) and records the same
   under "bhtree".

Usage:  python tests/golden/make_golden.py [--inputs-only]
"""
from __future__ import annotations

import hashlib
import json
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import cases  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref", "autogrid4")
REF_BH = os.path.join(ROOT, "oracle", "_ref", "autogrid4_bhtree")


def value_hash(path: str):
    data = open(path, "rb").read().split(b"\n", 6)
    body = data[6] if len(data) == 7 else b""
    return hashlib.sha256(body).hexdigest(), body.count(b"\n")


def main():
    for name in cases.SYN_CASES:
        cases.write_synthetic_inputs(name, os.path.join(cases.CASES_DIR, name))
    if "--inputs-only" in sys.argv:
        return
    if not os.path.exists(REF):
        sys.exit(f"{REF} missing: run `make -C oracle ref` first")
    out = {}
    for name in cases.ALL_CASES:
        with tempfile.TemporaryDirectory() as d:
            gpf = cases.materialize(name, d)
            p = subprocess.run([REF, "-p", gpf, "-l", "ref.glg"], cwd=d, capture_output=True)
            glg = open(os.path.join(d, "ref.glg"), errors="replace").read() if os.path.exists(
                os.path.join(d, "ref.glg")) else ""
            rec = {"rc": p.returncode, "success": "Successful Completion" in glg[-400:]}
            if rec["success"]:
                maps = {}
                for m in cases.map_names(os.path.join(d, gpf)):
                    h, n = value_hash(os.path.join(d, m))
                    maps[m] = {"sha256": h, "n": n,
                               "header": open(os.path.join(d, m)).read().split("\n")[:6]}
                rec["maps"] = maps
                for kind, fn in cases.distance_map_names(os.path.join(d, gpf)).items():
                    h, n = value_hash(os.path.join(d, fn))
                    rec[kind] = {"file": fn, "sha256": h, "n": n}
                fld = [ln for ln in open(os.path.join(d, f"{name.split('_CASE')[0] if name.startswith('x1hpv') else name}.maps.fld")).read().split("\n")
                       if not ln.startswith("# Created by")] if name in cases.SYN_CASES else None
                rec["fld"] = fld
                xyz = cases.xyz_name(os.path.join(d, gpf))
                rec["xyz"] = open(os.path.join(d, xyz)).read() if xyz and os.path.exists(os.path.join(d, xyz)) else None
                rec["glg"] = cases.glg_record(glg)
            out[name] = rec
            print(name, rec["rc"], rec["success"], len(rec.get("maps", {})))
    # the tree build (make -C oracle ref-bhtree) on the cases.BH_CASES variants
    if os.path.exists(REF_BH):
        bh = {}
        for name in cases.BH_CASES:
            with tempfile.TemporaryDirectory() as d:
                gpf = cases.materialize_bh(name, d)
                p = subprocess.run([REF_BH, "-p", gpf, "-l", "ref.glg"], cwd=d, capture_output=True)
                glg = open(os.path.join(d, "ref.glg"), errors="replace").read() if os.path.exists(
                    os.path.join(d, "ref.glg")) else ""
                rec = {"rc": p.returncode, "success": "Successful Completion" in glg[-400:],
                       "tree_line": [ln.strip() for ln in glg.split("\n") if "USE_BHTREE" in ln]}
                if rec["success"]:
                    rec["maps"] = {m: dict(zip(("sha256", "n"), value_hash(os.path.join(d, m))))
                                   for m in cases.map_names(os.path.join(d, gpf))}
                    for kind, fn in cases.distance_map_names(os.path.join(d, gpf)).items():
                        h, n = value_hash(os.path.join(d, fn))
                        rec[kind] = {"file": fn, "sha256": h, "n": n}
                bh[name] = rec
                print("bhtree", name, rec["rc"], rec["success"], rec["tree_line"])
        out["bhtree"] = bh
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
