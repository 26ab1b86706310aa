"""Shared parity helpers: tolerance check and case loading."""
from __future__ import annotations

import os

import numpy as np

import autogrid_amd as ag
import cases
import oracle_lib as ol

TOL_ABS = 1.0e-3   # kcal/mol
TOL_REL = 1.0e-3


def within(gpu: np.ndarray, ref: np.ndarray):
    """(ok, max_abs_err, rms_err, n_bad) under max(1e-3, 1e-3*|ref|) per point."""
    g = np.asarray(gpu, np.float64)
    r = np.asarray(ref, np.float64)
    err = np.abs(g - r)
    lim = np.maximum(TOL_ABS, TOL_REL * np.abs(r))
    bad = int(np.count_nonzero(~(err <= lim)))
    return bad == 0, float(err.max(initial=0.0)), float(np.sqrt(np.mean(err ** 2))) if err.size else 0.0, bad


def load_case_bh(name: str, workdir: str):
    """load_case for a tree-build case (cases.BH_CASES): the front end with
    the semantics of a -DUSE_BHTREE build."""
    gpf = cases.materialize_bh(name, workdir)
    cwd = os.getcwd()
    os.chdir(workdir)
    try:
        st = ag.Job(gpf, bhtree=True).prepare().state()
    finally:
        os.chdir(cwd)
    return st, gpf


def load_case(name: str, workdir: str):
    """(PreparedState from the engine's front end, oracle FP64 maps)."""
    gpf = cases.materialize(name, workdir)
    cwd = os.getcwd()
    os.chdir(workdir)
    try:
        st = ag.Job(gpf).prepare().state()
    finally:
        os.chdir(cwd)
    return st, gpf


def _printed_unit(s: str) -> float:
    """One unit in the last printed digit of a %6.2lf or %9.2le field."""
    if "e" in s:
        return 10.0 ** (int(s.split("e")[1]) - 2)
    return 10.0 ** -len(s.split(".")[1]) if "." in s else 1.0


def glg_mismatches(ours: dict, ref: dict) -> list:
    """Differences between two cases.glg_record() results.  Fixed text
    compares exactly; each printed extremum may differ from the reference's
    by the map tolerance max(1e-3, 1e-3 |v|) plus one unit of its last
    printed digit (both sides round an FP32 / FP64 value to 3 digits)."""
    out = []
    for k in ("calculating", "progress", "planes", "summary", "clamp_note", "completion"):
        if ours.get(k) != ref.get(k):
            out.append(f"{k}: {ours.get(k)!r} != {ref.get(k)!r}")
    if len(ours["extrema"]) != len(ref["extrema"]):
        out.append(f"extrema rows: {len(ours['extrema'])} != {len(ref['extrema'])}")
        return out
    for a, b in zip(ours["extrema"], ref["extrema"]):
        if a[0] != b[0] or a[1] != b[1] or a[4] != b[4]:
            out.append(f"extrema row labels: {a} != {b}")
            continue
        for j in (2, 3):
            va, vb = float(a[j]), float(b[j])
            lim = max(TOL_ABS, TOL_REL * abs(vb)) + _printed_unit(b[j]) * (1 + 1e-9)
            if abs(va - vb) > lim:
                out.append(f"extrema {b[1]} field {j}: {a[j]} vs {b[j]}")
    return out
