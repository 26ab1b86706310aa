"""The FP64 oracle against the reference autogrid4 and closed forms.

What pins the oracle:
  * golden.json holds SHA-256 hashes of the reference binary's map text for
    every case; the oracle's text must hash the same (byte-identical maps);
  * when oracle/_ref/autogrid4 is built here, it is also run live;
  * PyAutoDock maps from the reference's own tests;
  * closed forms for eps(r), sol_fn, the LJ known answer of
    Tests/hsg1_sm_nbpcoeffs.gpf, min-window smoothing and a one-atom e map.
"""
from __future__ import annotations

import dataclasses
import hashlib
import json
import os
import subprocess

import numpy as np
import pytest

import autogrid_amd as ag
import cases
import oracle_lib as ol
import parity

GOLDEN = json.load(open(os.path.join(cases.GOLDEN, "golden.json")))
SUCCESS = [n for n in cases.ALL_CASES if GOLDEN[n]["success"]]


@pytest.mark.parametrize("name", SUCCESS)
def test_oracle_text_equals_reference(name, tmp_path):
    st, gpf = parity.load_case(name, str(tmp_path))
    g = GOLDEN[name]
    if st.floating_grid or st.constriction_grid:
        m, fm, cm = ol.maps(ol.oracle_state(st), fmap=bool(st.floating_grid), cmap=bool(st.constriction_grid))
        if st.floating_grid:
            assert hashlib.sha256(ol.format_rmin_map(fm)).hexdigest() == g["fmap"]["sha256"]
        if st.constriction_grid:
            assert hashlib.sha256(ol.format_count_map(cm)).hexdigest() == g["cmap"]["sha256"]
    else:
        m = ol.maps(ol.oracle_state(st))
    for k, fn in enumerate(cases.map_names(os.path.join(tmp_path, gpf))):
        text = ol.format_map(m[k])
        assert hashlib.sha256(text).hexdigest() == g["maps"][fn]["sha256"], fn


BH_GOLDEN = GOLDEN.get("bhtree", {})
BH_SUCCESS = [n for n in cases.BH_CASES if BH_GOLDEN.get(n, {}).get("success")]


@pytest.mark.parametrize("name", BH_SUCCESS)
def test_oracle_tree_build_text_equals_reference(name, tmp_path):
    """The oracle in the tree build's mode (FP32 radius queries, the 1.5 A
    collision rule) against the reference built with -DUSE_BHTREE."""
    st, gpf = parity.load_case_bh(name, str(tmp_path))
    interior = "map_receptor_interior on" in cases.BH_CASES[name][1]
    assert st.bhtree == (2 if interior else 1)
    g = BH_GOLDEN[name]
    m, fm, cm = ol.maps(ol.oracle_state(st), fmap=True, cmap=True)
    if st.floating_grid:
        assert hashlib.sha256(ol.format_rmin_map(fm)).hexdigest() == g["fmap"]["sha256"]
    if st.constriction_grid:
        assert hashlib.sha256(ol.format_count_map(cm)).hexdigest() == g["cmap"]["sha256"]
    for k, fn in enumerate(cases.map_names(os.path.join(tmp_path, gpf))):
        text = ol.format_map(m[k])
        assert hashlib.sha256(text).hexdigest() == g["maps"][fn]["sha256"], fn


def test_tree_build_cases_differ_from_default():
    """The tree build's maps are not the default build's (collisions zero
    e, FP32 distances move bins), so the tree-build pins are not vacuous."""
    assert BH_SUCCESS, "no tree-build fixtures"
    differ = 0
    for name in BH_SUCCESS:
        base, extra = cases.BH_CASES[name]
        if extra:
            continue
        for fn, rec in BH_GOLDEN[name]["maps"].items():
            differ += rec["sha256"] != GOLDEN[base]["maps"][fn]["sha256"]
    assert differ > 0


def test_map_receptor_interior_off_needs_the_tree(tmp_path):
    """'map_receptor_interior off' is fatal without the tree build
    (main.cpp:1655-1656) and accepted with it."""
    gpf = cases.materialize_bh("interior_off", str(tmp_path))
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        with pytest.raises(ag.AgError, match="map_receptor_interior off"):
            ag.Job(gpf)
        assert ag.Job(gpf, bhtree=True).state().bhtree == 1
    finally:
        os.chdir(cwd)


@pytest.mark.skipif(not os.path.exists(ol.REF_BIN), reason="reference autogrid4 not built here")
@pytest.mark.parametrize("name", ["hsg1_sm", "disorder", "sepdesolv"])
def test_live_reference_run_matches_golden(name, tmp_path):
    gpf = cases.materialize(name, str(tmp_path))
    p = subprocess.run([ol.REF_BIN, "-p", gpf, "-l", "r.glg"], cwd=tmp_path, capture_output=True, timeout=120)
    assert p.returncode == 0
    for fn in cases.map_names(os.path.join(tmp_path, gpf)):
        body = open(tmp_path / fn, "rb").read().split(b"\n", 6)[6]
        assert hashlib.sha256(body).hexdigest() == GOLDEN[name]["maps"][fn]["sha256"]


@pytest.mark.parametrize("t", ["A", "C", "HD", "N", "NA", "OA", "e", "d"])
def test_oracle_against_pyautodock_fixture(t, tmp_path):
    """Tests/py_hsg1_sm.*.map (PyAutoDock, 125 points).  The reference's own
    check divides by a signed value; with the sign fixed the reference is
    within 0.003 of these maps, which is the bar here."""
    st, gpf = parity.load_case("hsg1_sm", str(tmp_path))
    m = ol.maps(ol.oracle_state(st))
    k = ["A", "C", "HD", "N", "NA", "OA", "e", "d"].index(t)
    mine = np.array([float(x) for x in ol.format_map(m[k]).decode().split("\n") if x])
    py = np.array([float(x) for x in open(os.path.join(cases.REFTESTS, f"py_hsg1_sm.{t}.map")).read().split("\n")[6:]
                   if x])
    assert mine.shape == py.shape == (125,)
    assert np.abs(mine - py).max() <= 0.005
    big = np.abs(mine) > 0.30   # the reference's own (signed) test passes as well
    assert np.all(((np.abs(mine - py) / mine) * 100.0 < 0.5)[big])
    assert np.all((np.abs(mine - py) < 0.01)[~big])


def test_dielectric_closed_form():
    r = np.arange(1, 8192) / 128.0
    A, lam, eps0, k = -8.5525, 0.003627, 78.4, 7.7839
    B = eps0 - A
    eps = A + B / (1.0 + k * np.exp(-lam * B * r))
    got = np.array([ol.olib().orc_ddd(float(x)) for x in r[::37]])
    assert np.allclose(got, eps[::37], rtol=1e-13, atol=0)
    reps = np.zeros(8192)
    ol.olib().orc_r_eps(reps.ctypes.data)
    assert reps[0] == 332.0 and np.allclose(reps[1:], 332.0 / eps, rtol=1e-13)


def test_sol_fn_closed_form():
    out = np.zeros(8192)
    ol.olib().orc_sol_fn(0.1322, out.ctypes.data)
    r = np.arange(8192) / 128.0
    ref = 0.1322 * np.exp(-r * r / (2 * 3.6 * 3.6))
    assert out[0] == 0.0 and np.allclose(out[1:], ref[1:], rtol=1e-14)


def _lj_state(tmp_path, smooth):
    st, _ = parity.load_case("hsg1_sm", str(tmp_path))
    pairs = st.pairs.copy()
    pairs["cA"], pairs["cB"], pairs["Rij"], pairs["eps"], pairs["xA"], pairs["xB"] = 0, 0, 1.5, 10.0, 12, 6
    return dataclasses.replace(st, pairs=pairs, r_smooth=smooth, _keep=[])


def test_lj_known_answer_and_minimum(tmp_path):
    """Tests/hsg1_sm_nbpcoeffs.gpf: r=1.5, eps=10, 12-6 gives cA=1297.46337890625,
    cB=227.8125; the unsmoothed table equals cA/r^12 - cB/r^6 and its minimum
    is -eps at r = 1.5 (bin 192)."""
    st = _lj_state(tmp_path, 0.0)
    ev, _, _ = ol.tables(st)
    r = np.arange(1, 1023) / 128.0
    ref = np.minimum(1e5, 1297.46337890625 / r ** 12 - 227.8125 / r ** 6).astype(np.float32)
    row = ev[0, 1]   # C map, first receptor type
    assert np.allclose(row[1:1023], ref, rtol=2e-7, atol=1e-6)
    assert row[0] == np.float32(1e5) and row[1023] == 0.0
    assert int(np.argmin(row)) == 192 and abs(row[192] + 10.0) < 1e-5


def test_min_window_smoothing_brute_force(tmp_path):
    st0 = _lj_state(tmp_path / "a", 0.0)
    st1 = _lj_state(tmp_path / "b", 0.5)
    raw, _, _ = ol.tables(st0)
    sm, _, _ = ol.tables(st1)
    row = raw[0, 1].astype(np.float64)
    half = 32   # floor(0.5 * 128 / 2)
    brute = np.array([min(100000.0, row[max(0, k - half):min(1024, k + half + 1)].min()) for k in range(1024)])
    assert np.array_equal(sm[0, 1], brute.astype(np.float32))


def test_single_atom_electrostatic_map(tmp_path):
    """One charged atom: e(p) = q*332*c/(max(r,0.5)*eps(bin/128))."""
    st, _ = parity.load_case("hsg1_sm", str(tmp_path))
    q = 0.437
    one = dataclasses.replace(st, coord=np.array([[0.1, -0.2, 0.05]]), charge=np.array([q]), vol=np.zeros(1),
                              solpar=np.zeros(1), type=np.zeros(1, np.int32), hbond=np.zeros(1, np.int32),
                              rexp=np.zeros(1, np.int32), rvector=np.zeros((1, 3)), rvector2=np.zeros((1, 3)),
                              disorder=np.zeros(1, np.uint8), maps=st.maps[:0], elec_index=0, dsolv_index=-1,
                              n_maps=1, e_vdw_hb=st.e_vdw_hb[:, :0], _keep=[])
    e = ol.maps(one)[0]
    nz, ny, nx = e.shape
    iz, iy, ix = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    c = np.stack([(ix - st.ne[0]) * st.spacing, (iy - st.ne[1]) * st.spacing, (iz - st.ne[2]) * st.spacing], -1)
    r = np.linalg.norm(c - one.coord[0], axis=-1)
    b = np.minimum((r * 128).astype(int), 8191)
    ref = q * 332.0 * st.coeff_estat / (np.maximum(r, 0.5) * np.where(b == 0, 1.0, [[[ol.olib().orc_ddd(k / 128.0)
                                                                                     for k in row] for row in pl]
                                                                                   for pl in b]))
    assert np.allclose(e, ref, rtol=1e-12)


def test_electrostatics_is_linear_in_charge(tmp_path):
    st, _ = parity.load_case("hsg1_sm", str(tmp_path))
    ost = ol.oracle_state(st)
    e1 = ol.maps(ost)[st.elec_index]
    e2 = ol.maps(dataclasses.replace(ost, charge=2.0 * ost.charge, _keep=[]))[st.elec_index]
    assert np.array_equal(2.0 * e1, e2)


def test_count_pairs_against_brute_force(tmp_path):
    """orc_count_pairs (W_A) equals an all-pairs count with numpy's hypot on
    the small case, plane ranges add up, and an atom exactly 8 A from a
    point counts (the reference skips only r > 8)."""
    import dataclasses
    st, _ = parity.load_case("small", str(tmp_path))
    h = st.spacing
    want = 0
    for iz in (0, 7, 20, 40):
        z = (iz - st.ne[2]) * h
        y = (np.arange(st.n1[1]) - st.ne[1]) * h
        x = (np.arange(st.n1[0]) - st.ne[0]) * h
        X, Y = np.meshgrid(x, y, indexing="xy")
        a = st.coord
        r = np.hypot(a[:, 0, None, None] - X, np.hypot(a[:, 1, None, None] - Y, a[:, 2, None, None] - z))
        n = int(np.count_nonzero(~(r > 8.0)))
        assert ol.count_pairs(st, iz, iz + 1) == n, iz
        want += n
    assert want > 0
    total = ol.count_pairs(st)
    assert total == ol.count_pairs(st, 0, 20) + ol.count_pairs(st, 20, st.n1[2])
    # an atom at exactly 8.0 A along x from the grid origin point counts
    c = np.array([[8.0, 0.0, 0.0]])
    one = dataclasses.replace(st, coord=c, charge=np.zeros(1), vol=np.zeros(1), solpar=np.zeros(1),
                              type=st.type[:1].copy(), hbond=np.zeros(1, np.int32), rexp=np.zeros(1, np.int32),
                              rvector=np.zeros((1, 3)), rvector2=np.zeros((1, 3)), disorder=np.zeros(1, np.uint8),
                              _keep=[])
    iz = st.ne[2]
    X = (np.arange(st.n1[0]) - st.ne[0]) * h
    Y = (np.arange(st.n1[1]) - st.ne[1]) * h
    r = np.hypot(8.0 - X[None, :], np.hypot(0.0 - Y[:, None], 0.0))
    assert ol.count_pairs(one, iz, iz + 1) == int(np.count_nonzero(r <= 8.0))
    assert np.any(r == 8.0)
