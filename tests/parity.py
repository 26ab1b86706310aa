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
