"""Device parity: the HIP map path against the FP64 oracle on every case.

The oracle reproduces the reference autogrid4 byte for byte on these inputs
(tests/test_oracle.py), so agreement with it is agreement with the
reference.  Bar: |gpu - oracle| <= max(1e-3 kcal/mol, 1e-3 * |oracle|) at
every point of every map.
"""
from __future__ import annotations

import json
import os
import subprocess

import numpy as np
import pytest

import autogrid_amd as ag
import cases
import oracle_lib as ol
import parity

pytestmark = pytest.mark.gpu

GOLDEN = json.load(open(os.path.join(cases.GOLDEN, "golden.json")))
SUCCESS = [n for n in cases.ALL_CASES if GOLDEN[n]["success"]]


@pytest.fixture(scope="module")
def ctx():
    c = ag.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("name", SUCCESS)
def test_maps_match_oracle(ctx, name, tmp_path):
    print(f"[gpu] case {name}", flush=True)
    st, _ = parity.load_case(name, str(tmp_path))
    want_f, want_c = bool(st.floating_grid), bool(st.constriction_grid)
    if want_f or want_c:
        ref, rfm, rcm = ol.maps(ol.oracle_state(st), fmap=want_f, cmap=want_c)
    else:
        ref = ol.maps(ol.oracle_state(st))
    got = ctx.maps_plain(st)
    assert got.shape == ref.shape
    if want_f:   # distance to the nearest atom, A
        assert np.abs(ctx.last_fmap.astype(np.float64) - rfm).max() <= 1e-4
    if want_c:   # atom counts are exact
        assert np.array_equal(ctx.last_cmap, rcm)
    report = []
    for k in range(st.n_maps):
        ok, mx, rms, bad = parity.within(got[k], ref[k])
        report.append((k, ok, mx, rms, bad))
    print(name, [(k, f"{mx:.2e}", f"{rms:.2e}", bad) for k, ok, mx, rms, bad in report], flush=True)
    failed = [r for r in report if not r[1]]
    assert not failed, f"{name}: maps over tolerance {failed}"


BH_SUCCESS = [n for n in cases.BH_CASES if GOLDEN.get("bhtree", {}).get(n, {}).get("success")]


@pytest.mark.parametrize("name", BH_SUCCESS)
def test_tree_build_maps_match_oracle(ctx, name, tmp_path):
    """The -DUSE_BHTREE semantics (FP32 radius queries for Loop 2 and the
    constriction count, the 1.5 A collision rule) against the oracle, which
    reproduces the reference's tree build byte for byte."""
    print(f"[gpu] tree-build case {name}", flush=True)
    st, _ = parity.load_case_bh(name, str(tmp_path))
    assert st.bhtree in (1, 2)
    want_f, want_c = bool(st.floating_grid), bool(st.constriction_grid)
    ref, rfm, rcm = ol.maps(ol.oracle_state(st), fmap=True, cmap=True)
    got = ctx.maps_plain(st)
    if want_f:
        assert np.abs(ctx.last_fmap.astype(np.float64) - rfm).max() <= 1e-4
    if want_c:
        assert np.array_equal(ctx.last_cmap, rcm)
    if st.elec_index >= 0:
        # collision points hold exactly 0 on both sides
        zero = rfm == 1.5 if st.bhtree == 1 else np.zeros_like(rfm, bool)
        assert np.all(got[st.elec_index][zero] == 0.0)
    report = []
    for k in range(st.n_maps):
        ok, mx, rms, bad = parity.within(got[k], ref[k])
        report.append((k, ok, mx, rms, bad))
    print(name, [(k, f"{mx:.2e}", f"{rms:.2e}", bad) for k, ok, mx, rms, bad in report], flush=True)
    failed = [r for r in report if not r[1]]
    assert not failed, f"{name}: maps over tolerance {failed}"


def test_repeat_is_bitwise_identical(ctx, tmp_path):
    st, _ = parity.load_case("small", str(tmp_path))
    a = ctx.maps_plain(st)
    b = ctx.maps_plain(st)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("name", ["small", "x1hpv_CASE0"])
def test_near_kernel_forms_agree(ctx, name, tmp_path):
    """A map set without a constriction grid, tree-build queries or
    disordered hydroxyls runs the near kernel's plain form (those branches
    compiled out); asking for a constriction grid runs the general form.
    The maps they share must be the same bits: the forms differ in branches
    only, not in arithmetic."""
    import dataclasses
    st, _ = parity.load_case(name, str(tmp_path))
    assert not st.constriction_grid and not st.bhtree and not np.any(st.disorder)
    plain = ctx.maps_plain(st)
    general = ctx.maps_plain(dataclasses.replace(st, constriction_grid=1, constriction_cutoff=6.5, _keep=[]))
    assert ctx.last_cmap is not None and np.any(ctx.last_cmap > 0)
    assert np.array_equal(plain.view(np.uint32), general.view(np.uint32))


def test_context_reuse_across_sizes(ctx, tmp_path):
    """Scratch grows and is reused: big, small, big again give the same maps."""
    big, _ = parity.load_case("small", str(tmp_path / "a"))
    small, _ = parity.load_case("hsg1_sm", str(tmp_path / "b"))
    m1 = ctx.maps_plain(big)
    m2 = ctx.maps_plain(small)
    m3 = ctx.maps_plain(big)
    assert np.array_equal(m1, m3)
    ok, *_ = parity.within(m2, ol.maps(ol.oracle_state(small)))
    assert ok


def test_device_outputs(ctx, tmp_path):
    """Outputs written straight into device memory equal the host path."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so.7")   # the runtime the engine is bound to
    st, _ = parity.load_case("hsg1_sm", str(tmp_path))
    host = ctx.maps_plain(st)
    ptr = C.c_void_p()
    assert hip.hipMalloc(C.byref(ptr), C.c_size_t(host.nbytes)) == 0
    try:
        ctx.maps_plain_device(st, ptr.value)
        back = np.empty_like(host)
        assert hip.hipMemcpy(C.c_void_p(back.ctypes.data), ptr, C.c_size_t(host.nbytes), 2) == 0  # D2H
    finally:
        hip.hipFree(ptr)
    assert np.array_equal(back, host)


@pytest.mark.parametrize("name", ["hsg1_sm", "x1hpv_CASE0", "covalent", "sepdesolv", "fmap_cmap", "small", "all13",
                                  "nodiel", "constdiel"])
def test_cli_dropin(name, tmp_path):
    """autogrid4 drop-in: exit 0, the log lines consumers parse (per-plane
    table, extrema table, clamp note, completion), the .xyz text, reference
    headers and values."""
    print(f"[gpu] cli {name}", flush=True)
    gpf = cases.materialize(name, str(tmp_path))
    p = subprocess.run([ag.CLI_PATH, "-p", gpf, "-l", "out.glg"], cwd=tmp_path, capture_output=True,
                       timeout=120)
    assert p.returncode == 0, p.stderr.decode()
    glg = open(tmp_path / "out.glg").read()
    assert "Successful Completion" in "".join(glg.splitlines(True)[-10:])
    st, _ = parity.load_case(name, str(tmp_path / "oracle"))
    ref = ol.maps(ol.oracle_state(st))
    g = GOLDEN[name]
    # the log lines consumers parse: fixed text exact, extrema to tolerance
    bad = parity.glg_mismatches(cases.glg_record(glg), g["glg"])
    assert not bad, bad
    assert open(tmp_path / cases.xyz_name(os.path.join(tmp_path, gpf))).read() == g["xyz"]
    for k, m in enumerate(cases.map_names(os.path.join(tmp_path, gpf))):
        lines = open(tmp_path / m).read().split("\n")
        assert lines[:6] == g["maps"][m]["header"]
        vals = np.array([float(x) for x in lines[6:] if x])
        assert vals.size == g["maps"][m]["n"]
        # the text carries 3 decimals: allow the rounding of both sides
        err = np.abs(vals - ref[k].reshape(-1))
        lim = np.maximum(2e-3, 1e-3 * np.abs(ref[k].reshape(-1)) + 5e-4)
        assert np.all(err <= lim), f"{m}: max {err.max()}"
    if g.get("fld"):
        fld = [ln for ln in open(tmp_path / f"{name}.maps.fld").read().split("\n") if not ln.startswith("# Created by")]
        assert fld == g["fld"]
    if "cmap" in g:   # integer counts: byte-identical text
        body = open(tmp_path / g["cmap"]["file"], "rb").read().split(b"\n", 6)[6]
        import hashlib
        assert hashlib.sha256(body).hexdigest() == g["cmap"]["sha256"]
    if "fmap" in g:
        vals = np.array([float(x) for x in open(tmp_path / g["fmap"]["file"]).read().split("\n")[6:] if x])
        assert vals.size == g["fmap"]["n"]


@pytest.mark.parametrize("name", ["fmap_cmap", "interior_off"])
def test_cli_dropin_tree_build(name, tmp_path):
    """autogrid4 --bhtree: the tree build's outputs.  The constriction counts
    are byte-identical to the reference tree build's text; the maps hold the
    oracle's values within the text's rounding."""
    import hashlib
    print(f"[gpu] cli --bhtree {name}", flush=True)
    gpf = cases.materialize_bh(name, str(tmp_path))
    p = subprocess.run([ag.CLI_PATH, "--bhtree", "-p", gpf, "-l", "out.glg"], cwd=tmp_path, capture_output=True,
                       timeout=120)
    assert p.returncode == 0, p.stderr.decode()
    glg = open(tmp_path / "out.glg").read()
    assert "Successful Completion" in "".join(glg.splitlines(True)[-10:])
    st, _ = parity.load_case_bh(name, str(tmp_path / "oracle"))
    ref = ol.maps(ol.oracle_state(st))
    g = GOLDEN["bhtree"][name]
    for k, m in enumerate(cases.map_names(os.path.join(tmp_path, gpf))):
        vals = np.array([float(x) for x in open(tmp_path / m).read().split("\n")[6:] if x])
        assert vals.size == g["maps"][m]["n"]
        err = np.abs(vals - ref[k].reshape(-1))
        lim = np.maximum(2e-3, 1e-3 * np.abs(ref[k].reshape(-1)) + 5e-4)
        assert np.all(err <= lim), f"{m}: max {err.max()}"
    if "cmap" in g:
        body = open(tmp_path / g["cmap"]["file"], "rb").read().split(b"\n", 6)[6]
        assert hashlib.sha256(body).hexdigest() == g["cmap"]["sha256"]


def test_cli_binary_maps(ctx, tmp_path):
    """autogrid4 --binary-maps: the .bin files hold the very floats the ASCII
    text was formatted from, and those equal the library's map set bit for
    bit (same kernels, fixed-order sums)."""
    print("[gpu] cli --binary-maps", flush=True)
    name = "fmap_cmap"
    gpf = cases.materialize(name, str(tmp_path))
    p = subprocess.run([ag.CLI_PATH, "--binary-maps", "-p", gpf, "-l", "out.glg"], cwd=tmp_path,
                       capture_output=True, timeout=120)
    assert p.returncode == 0, p.stderr.decode()
    st, _ = parity.load_case(name, str(tmp_path / "lib"))
    want = ctx.maps_plain(st)
    for k, fn in enumerate(cases.map_names(os.path.join(tmp_path, gpf))):
        b = ag.read_binary_map(str(tmp_path / (fn + ".bin")), st.n1)
        assert np.array_equal(b, want[k]), fn
        text = open(tmp_path / fn, "rb").read().split(b"\n", 6)[6]
        assert text == ag.format_map_values(b)
    g = GOLDEN[name]
    cm = np.fromfile(tmp_path / (g["cmap"]["file"] + ".bin"), "<f4")
    assert np.array_equal(cm, ctx.last_cmap.reshape(-1).astype(np.float32))
    assert (tmp_path / f"{name}.maps.bin.fld").exists()


@pytest.mark.parametrize("npts", ["2 2 2", "2 4 6", "2 2 40", "10 1 3", "40 2 2"])
def test_small_and_ragged_grids(ctx, npts, tmp_path):
    """Grids smaller than a brick or a far-kernel run, and ragged ones (odd
    npts are rounded as check_size.cpp does): the small receptor's maps on
    each against the oracle, at every point."""
    print(f"[gpu] grid npts {npts}", flush=True)
    gpf = cases.materialize("small", str(tmp_path))
    path = os.path.join(tmp_path, gpf)
    lines = open(path).read().split("\n")
    lines = [f"npts {npts}" if ln.startswith("npts") else ln for ln in lines]
    open(path, "w").write("\n".join(lines))
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        st = ag.Job(gpf).prepare().state()
    finally:
        os.chdir(cwd)
    ref = ol.maps(ol.oracle_state(st))
    got = ctx.maps_plain(st)
    assert got.shape == ref.shape
    for k in range(st.n_maps):
        ok, mx, rms, bad = parity.within(got[k], ref[k])
        assert ok, f"npts {npts} map {k}: max {mx} rms {rms} bad {bad}"


@pytest.mark.parametrize("n_atoms", [1, 63, 64, 65, 129])
def test_candidate_counts_around_index_blocks(ctx, n_atoms, tmp_path):
    """Every brick of an 11^3 grid sees the same n_atoms candidates: the
    atoms are packed into a 3 A ball at the grid centre, within 8 A of every
    point.  The near kernel reads candidate indices 64 at a time and loads
    each record one candidate ahead, so these counts put the last candidate
    of a block, the first of the next and a lone candidate through the loop.
    Maps against the oracle at every point."""
    import dataclasses
    print(f"[gpu] {n_atoms} candidates per brick", flush=True)
    gpf = cases.materialize("small", str(tmp_path))
    path = os.path.join(tmp_path, gpf)
    lines = open(path).read().split("\n")
    lines = ["npts 10 10 10" if ln.startswith("npts") else ln for ln in lines]
    open(path, "w").write("\n".join(lines))
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        st = ag.Job(gpf).prepare().state()
    finally:
        os.chdir(cwd)
    # the reference needs a donor H among the atoms (its closest-H search)
    donor = int(np.flatnonzero(st.hbond == 2)[0])
    pick = np.concatenate([[donor], np.setdiff1d(np.arange(n_atoms + 1), [donor])[:n_atoms - 1]])
    assert len(pick) == n_atoms and len(set(pick.tolist())) == n_atoms
    c = st.coord[pick] - st.coord[pick].mean(axis=0)
    scale = 3.0 / max(float(np.linalg.norm(c, axis=1).max()), 1e-9)
    st = dataclasses.replace(
        st, coord=np.ascontiguousarray(c * scale), charge=st.charge[pick].copy(), vol=st.vol[pick].copy(),
        solpar=st.solpar[pick].copy(), type=st.type[pick].copy(), hbond=st.hbond[pick].copy(),
        rexp=st.rexp[pick].copy(), rvector=st.rvector[pick].copy(), rvector2=st.rvector2[pick].copy(),
        disorder=st.disorder[pick].copy(), _keep=[])
    # bond vectors for the packed geometry (the oracle derives its own from it)
    rexp, rv, rv2, dis = ag.bond_vectors(st, st.coord, st.type, st.hbond)
    st = dataclasses.replace(st, rexp=rexp, rvector=rv, rvector2=rv2, disorder=dis, _keep=[])
    # every point lies within 8 A of every atom, so each brick stages them all
    half = 5 * st.spacing * np.sqrt(3.0)
    assert half + 3.0 < 8.0
    ref = ol.maps(ol.oracle_state(st))
    got = ctx.maps_plain(st)
    assert got.shape == ref.shape
    for k in range(st.n_maps):
        ok, mx, rms, bad = parity.within(got[k], ref[k])
        assert ok, f"{n_atoms} atoms map {k}: max {mx} rms {rms} bad {bad}"


def test_closest_donor_h_ties_take_the_lower_index(ctx, tmp_path):
    """Two C-H donor groups whose H sit at mirror positions about the plane x = 0:
    every grid point in that plane is exactly equidistant from both H, and
    the reference keeps the lower atom index as the closest H.  The parents
    point in different directions, so the choice changes the H-bonding maps
    there (the other H gets the angular ramp).  Maps against the oracle."""
    import dataclasses
    gpf = cases.materialize("small", str(tmp_path))
    path = os.path.join(tmp_path, gpf)
    lines = open(path).read().split("\n")
    lines = ["npts 10 10 10" if ln.startswith("npts") else ln for ln in lines]
    open(path, "w").write("\n".join(lines))
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        st = ag.Job(gpf).prepare().state()
    finally:
        os.chdir(cwd)
    tn = st.t_types
    donor = int(np.flatnonzero(st.hbond == 2)[0])
    parent = int(np.flatnonzero(st.type == tn["t_carbon"])[0])
    pick = np.array([parent, donor, parent, donor])
    coord = np.array([[1.75, 0.2, 0.1], [0.75, 0.2, 0.1],       # C1 - H1, bond along -x
                      [-0.75, 1.2, 0.1], [-0.75, 0.2, 0.1]])    # C2 - H2, bond along -y
    st = dataclasses.replace(
        st, coord=coord, charge=st.charge[pick].copy(), vol=st.vol[pick].copy(), solpar=st.solpar[pick].copy(),
        type=st.type[pick].copy(), hbond=st.hbond[pick].copy(), rexp=st.rexp[pick].copy(),
        rvector=st.rvector[pick].copy(), rvector2=st.rvector2[pick].copy(), disorder=st.disorder[pick].copy(),
        _keep=[])
    rexp, rv, rv2, dis = ag.bond_vectors(st, st.coord, st.type, st.hbond)
    st = dataclasses.replace(st, rexp=rexp, rvector=rv, rvector2=rv2, disorder=dis, _keep=[])
    assert np.allclose(rv[1], [-1, 0, 0]) and np.allclose(rv[3], [0, -1, 0])
    ref = ol.maps(ol.oracle_state(st))
    got = ctx.maps_plain(st)
    for k in range(st.n_maps):
        ok, mx, rms, bad = parity.within(got[k], ref[k])
        assert ok, f"map {k}: max {mx} rms {rms} bad {bad}"


def test_overfull_brick_runs_paged_against_oracle(ctx, tmp_path):
    """The near kernel stages at most 2,048 candidate atoms and 256 donor H
    per 8x8x4 brick in LDS, about ten times protein density over the
    brick's 8.03 A staging volume.  A denser receptor sets the status word;
    the plan then switches to the paged near kernel, which stages and
    processes the lists page by page in the same order, and runs the map
    set again.  6,000 atoms (about 360 donor H) in a 6 A ball: every map
    against the oracle at every point; the same context then computes an
    ordinary map set; and a batch holding both receptors (its plan switches
    as a whole) equals each one's own plain run bit for bit."""
    import dataclasses
    print("[gpu] overfull brick, paged near kernel", flush=True)
    gpf = cases.materialize("small", str(tmp_path))
    path = os.path.join(tmp_path, gpf)
    lines = open(path).read().split("\n")
    lines = ["npts 10 10 10" if ln.startswith("npts") else ln for ln in lines]
    open(path, "w").write("\n".join(lines))
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        base = ag.Job(gpf).prepare().state()
    finally:
        os.chdir(cwd)
    # 6,000 atoms: the receptor's 1,000 six times, packed into a 6 A ball
    pick = np.tile(np.arange(len(base.coord)), 6)
    rng = np.random.default_rng(11)
    u = rng.normal(size=(len(pick), 3))
    u *= (6.0 * rng.random(len(pick)) ** (1 / 3) / np.linalg.norm(u, axis=1))[:, None]
    st = dataclasses.replace(
        base, coord=np.ascontiguousarray(u), charge=base.charge[pick].copy(), vol=base.vol[pick].copy(),
        solpar=base.solpar[pick].copy(), type=base.type[pick].copy(), hbond=base.hbond[pick].copy(),
        rexp=base.rexp[pick].copy(), rvector=base.rvector[pick].copy(), rvector2=base.rvector2[pick].copy(),
        disorder=base.disorder[pick].copy(), _keep=[])
    rexp, rv, rv2, dis = ag.bond_vectors(st, st.coord, st.type, st.hbond)
    st = dataclasses.replace(st, rexp=rexp, rvector=rv, rvector2=rv2, disorder=dis, _keep=[])
    assert (st.hbond == 2).sum() > 256
    got = ctx.maps_plain(st)
    ref = ol.maps(ol.oracle_state(st))
    for k in range(st.n_maps):
        ok, mx, rms, bad = parity.within(got[k], ref[k])
        assert ok, f"dense map {k}: max {mx} rms {rms} bad {bad}"
    plain = ctx.maps_plain(base)
    z = base.n1[2] // 2
    bref = ol.maps(ol.oracle_state(base), z, z + 1)[:, z]
    for k in range(base.n_maps):
        ok, mx, rms, bad = parity.within(plain[k, z], bref[k])
        assert ok, f"map {k}: max {mx} rms {rms} bad {bad}"
    both = ctx.maps_batch([base, st], tune={"cell_size": 4.0, "cube": True})
    assert np.array_equal(both[0].view(np.uint32), plain.view(np.uint32))
    assert np.array_equal(both[1].view(np.uint32), got.view(np.uint32))


def test_protein_density_box_against_oracle(ctx, tmp_path):
    """The reference's x1hpv receptor (1,849 atoms, a real protein's
    density) with its CASE0 map set on a 121^3 box, the bench's protein
    workload: many bricks stage over 100 candidates, and the bricks go to
    the XCDs in chunks of 16 over a grid whose work per brick varies.  Three
    planes of every map against the oracle, to the parity bar."""
    print("[gpu] x1hpv 121^3 planes vs oracle", flush=True)
    gpf = cases.materialize("x1hpv_CASE0", str(tmp_path))
    path = os.path.join(tmp_path, gpf)
    text = open(path).read()
    assert "npts 6 6 6" in text
    open(path, "w").write(text.replace("npts 6 6 6", "npts 120 120 120"))
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        st = ag.Job(gpf).prepare().state()
    finally:
        os.chdir(cwd)
    assert st.n1 == (121, 121, 121) and st.n_atoms == 1849
    got = ctx.maps_plain(st)
    ost = ol.oracle_state(st)
    for z in (0, 60, 120):
        ref = ol.maps(ost, z, z + 1)[:, z]
        for k in range(st.n_maps):
            ok, mx, rms, bad = parity.within(got[k, z], ref[k])
            assert ok, f"plane {z} map {k}: max {mx} rms {rms} bad {bad}"
