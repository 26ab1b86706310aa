"""Host side of the batched path: the precompute entry points and the
ensemble generator (no GPU)."""
from __future__ import annotations

import os

import numpy as np

import autogrid_amd as ag
import cases
import oracle_lib as ol
from autogrid_amd import synth


def test_build_tables_and_bond_vectors_match_job(tmp_path):
    """ag_build_tables / ag_bond_vectors give the bits ag_job_prepare gives,
    which the oracle tests pin to the reference."""
    gpf = cases.materialize("disorder", str(tmp_path))
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        st = ag.Job(gpf).prepare().state()
    finally:
        os.chdir(cwd)
    ev, sol, reps = ag.build_tables(st)
    assert np.array_equal(ev.view(np.uint32), st.e_vdw_hb.view(np.uint32))
    assert np.array_equal(sol, st.sol_fn) and np.array_equal(reps, st.r_eps)
    rexp, rv, rv2, dis = ag.bond_vectors(st, st.coord, st.type, st.hbond)
    assert np.array_equal(rexp, st.rexp) and np.array_equal(rv, st.rvector)
    assert np.array_equal(rv2, st.rvector2) and np.array_equal(dis, st.disorder)
    # and the oracle agrees
    orexp, orv, orv2, odis = ol.bond_vectors(st)
    assert np.array_equal(orexp, rexp) and np.array_equal(orv, rv) and np.array_equal(odis, dis)


def test_ensemble_is_seeded_and_in_range(tmp_path):
    a = synth.make_ensemble(str(tmp_path / "a"), n_conf=8, seed=3001, base_atoms=3000, n_range=(600, 3000),
                            npts_range=(16, 40))
    b = synth.make_ensemble(str(tmp_path / "b"), n_conf=8, seed=3001, base_atoms=3000, n_range=(600, 3000),
                            npts_range=(16, 40))
    assert a.rec_types == b.rec_types and a.lig_types == ["A", "C", "HD", "N", "NA", "OA"]
    for ca, cb in zip(a.confs, b.confs):
        assert ca["n_atoms"] == cb["n_atoms"] and ca["npts"] == cb["npts"]
        assert np.array_equal(ca["coord"], cb["coord"])
        assert 600 <= ca["n_atoms"] <= 3000 and 16 <= ca["npts"] <= 40 and ca["npts"] % 2 == 0
    # no conformation ends on a donor hydrogen cut from its parent
    for c in a.confs:
        n = c["n_atoms"]
        assert n == a.base.n_atoms or a.base.hbond[n] != 2
    st = synth.conformation_state(a, 0)
    assert st.n1 == (a.confs[0]["npts"] + 1,) * 3
    assert np.allclose(st.coord + a.confs[0]["center"], a.confs[0]["coord"])
    # hydrogens ride with their parents: the N/O-H bond length is unchanged by the jitter
    base_abs = a.base.coord + a.center
    h = np.nonzero(a.base.hbond[:st.n_atoms] == 2)[0]
    h = h[h > 0]
    d0 = np.linalg.norm(base_abs[h] - base_abs[h - 1], axis=1)
    d1 = np.linalg.norm(a.confs[0]["coord"][h] - a.confs[0]["coord"][h - 1], axis=1)
    moved_apart = np.abs(d1 - d0) > 1e-9
    # a loop boundary may fall between a parent and its hydrogen; elsewhere the bond is rigid
    assert moved_apart.sum() <= 6


def test_bond_vectors_batch_matches_single(tmp_path):
    ens = synth.make_ensemble(str(tmp_path), n_conf=5, seed=3002, base_atoms=2000, n_range=(300, 2000),
                              npts_range=(16, 24))
    raw = [synth.conformation_state(ens, i, bonds=False) for i in range(5)]
    got = ag.bond_vectors_batch(raw, n_threads=3)
    for st, g in zip(raw, got):
        rexp, rv, rv2, dis = ag.bond_vectors(st, st.coord, st.type, st.hbond)
        assert np.array_equal(g.rexp, rexp) and np.array_equal(g.rvector, rv)
        assert np.array_equal(g.rvector2, rv2) and np.array_equal(g.disorder, dis)


def test_maps_batch_rejects_bad_output_arrays(tmp_path):
    """Caller-supplied batch outputs are checked for dtype, shape and
    contiguity before any address reaches the library (no device call:
    the checks run first)."""
    import dataclasses

    import pytest
    ens = synth.make_ensemble(str(tmp_path), n_conf=2, seed=3001, base_atoms=800, n_range=(300, 600),
                              npts_range=(8, 12))
    states = [synth.conformation_state(ens, i) for i in range(2)]
    ctx = object.__new__(ag.Context)      # no device: validation precedes the library call
    good = [np.zeros((st.n_maps, st.n1[2], st.n1[1], st.n1[0]), np.float32) for st in states]
    bad_cases = [
        good[:1],                                                    # one array for two conformations
        [good[0], good[1].astype(np.float64)],                       # wrong dtype
        [good[0], good[1][:, :-1]],                                  # wrong shape
        [good[0], np.asfortranarray(good[1])],                       # not C-contiguous
        [good[1], good[0]] if states[0].points != states[1].points else [good[0], good[1][..., :1]],
    ]
    for outs in bad_cases:
        with pytest.raises(ValueError):
            ctx.maps_batch(states, outs=outs)
    sts = [dataclasses.replace(s, floating_grid=1, constriction_grid=1, _keep=[]) for s in states]
    fm = [np.empty(st.n1[::-1], np.float32) for st in sts]
    with pytest.raises(ValueError):          # cmap must be int32
        ctx.maps_batch(sts, outs=good, fmaps=fm, cmaps=[a.copy() for a in fm])
    with pytest.raises(ValueError):          # fmap of the wrong shape
        ctx.maps_batch(sts, outs=good, fmaps=[fm[0], fm[0][:1]], cmaps=[a.view(np.int32) for a in fm])
