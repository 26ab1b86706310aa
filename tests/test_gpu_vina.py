"""Vina-potential map sets on the GPU against the FP64 oracle.

Tables come from the engine's front end (ag_job_prepare_vina), which the CPU
tests pin bit for bit to the oracle's tables; the per-point sums are then
compared under max(1e-3 kcal/mol, 1e-3 |oracle|) at every point.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

import autogrid_amd as ag
import cases
import oracle_lib as ol
import parity
from autogrid_amd import synth

pytestmark = pytest.mark.gpu


def _vina_state(name, d):
    gpf = cases.materialize(name, str(d))
    cwd = os.getcwd()
    os.chdir(d)
    try:
        return ag.Job(gpf).prepare_vina().state()
    finally:
        os.chdir(cwd)


@pytest.mark.parametrize("name", ["small", "all13", "covalent", "disorder", "fmap_cmap"])
def test_vina_maps_against_oracle(name, tmp_path):
    print(f"[gpu] vina {name}", flush=True)
    st = _vina_state(name, tmp_path)
    with ag.Context(0) as ctx:
        got = ctx.maps_vina(st)
        fm, cm = ctx.last_fmap, ctx.last_cmap
        again = ctx.maps_vina(st)
    assert np.array_equal(got, again)
    if st.floating_grid or st.constriction_grid:
        ref, rfm, rcm = ol.maps_vina(st, fmap=bool(st.floating_grid), cmap=bool(st.constriction_grid))
        if fm is not None:
            assert parity.within(fm, rfm)[0]
        if cm is not None:
            assert np.array_equal(cm, rcm)
    else:
        ref = ol.maps_vina(st)
    for k in range(st.n_maps):
        ok, mx, rms, bad = parity.within(got[k], ref[k])
        assert ok, f"{name} vina map {k}: max {mx} rms {rms}, {bad} points over tolerance"
    for k in (st.elec_index, st.dsolv_index):
        if k >= 0:
            assert np.all(got[k] == 0)


def test_vina_full_workload_planes(tmp_path):
    """The 5,000-atom, 127^3-point map set with Vina tables, selected planes."""
    print("[gpu] vina full", flush=True)
    info = synth.make_workload("full", str(tmp_path))
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        st = ag.Job(info["gpf"]).prepare_vina().state()
    finally:
        os.chdir(cwd)
    with ag.Context(0) as ctx:
        got = ctx.maps_vina(st)
    worst = 0.0
    for z in (0, 63, 126):
        ref = ol.maps_vina(st, z, z + 1)[:, z]
        for k in range(st.n_maps):
            ok, mx, rms, bad = parity.within(got[k, z], ref[k])
            worst = max(worst, mx)
            assert ok, f"vina full map {k} plane {z}: max {mx} rms {rms} over tolerance {bad}"
    print(f"vina full: max|err| {worst:.3e}", flush=True)
