"""The library's write path (ag_job_write_maps) and the binary side-output,
on the host only: map files, .fld, headers and values from given maps.

* .map headers and the .fld text equal the reference's (golden.json);
* map values are the value text of main.cpp:3486-3497 for the given floats;
* with AG_JOB_BINARY_MAPS every map, the floating grid and the constriction
  counts are also written as raw little-endian FP32 (<file>.bin), and
  <name>.maps.bin.fld is the .fld with its variables pointing at them.
"""
from __future__ import annotations

import json
import os

import numpy as np

import autogrid_amd as ag
import cases

GOLDEN = json.load(open(os.path.join(cases.GOLDEN, "golden.json")))


def _job(tmp_path, name, **kw):
    gpf = cases.materialize(name, str(tmp_path))
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        job = ag.Job(gpf, "out.glg", files=True, **kw).prepare()
        return job, gpf
    finally:
        os.chdir(cwd)


def test_write_maps_ascii_and_binary(tmp_path):
    name = "fmap_cmap"
    job, gpf = _job(tmp_path, name, binary_maps=True)
    st = job.state()
    P = st.points
    rng = np.random.default_rng(4)
    maps = rng.normal(0, 3.0, size=(st.n_maps, P)).astype(np.float32)
    maps[0, :10] = [0.0, 5e-5, -5e-5, 1e-4, 0.0005, -0.0005, 123456.789, -0.0004999, 2.5e7, 1e5]
    fmap = rng.uniform(0.5, 6.0, size=P).astype(np.float32)
    cmap = rng.integers(0, 40, size=P).astype(np.int32)
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        job.write_maps(maps, fmap, cmap)
    finally:
        os.chdir(cwd)
    job.close()
    g = GOLDEN[name]
    names = cases.map_names(os.path.join(tmp_path, gpf))
    for k, fn in enumerate(names):
        lines = open(tmp_path / fn).read().split("\n")
        assert lines[:6] == g["maps"][fn]["header"]
        text = "\n".join(lines[6:]).encode()
        assert text == ag.format_map_values(maps[k])
        assert np.array_equal(ag.read_binary_map(str(tmp_path / (fn + ".bin")), st.n1).reshape(-1), maps[k])
    fm = np.fromfile(tmp_path / (g["fmap"]["file"] + ".bin"), "<f4")
    assert np.array_equal(fm, fmap)
    cm = np.fromfile(tmp_path / (g["cmap"]["file"] + ".bin"), "<f4")
    assert np.array_equal(cm, cmap.astype(np.float32))
    fld = [ln for ln in open(tmp_path / f"{name}.maps.fld").read().split("\n") if not ln.startswith("# Created by")]
    assert fld == g["fld"]
    bfld = [ln for ln in open(tmp_path / f"{name}.maps.bin.fld").read().split("\n")
            if not ln.startswith("# Created by")]
    assert len(bfld) == len(fld)
    nvar = 0
    for a, b in zip(fld, bfld):
        if a.startswith("variable"):
            k, f = a.split()[1], a.split()[2]
            assert b == f"variable {k} {f}.bin filetype=binary skip=0"
            nvar += 1
        else:
            assert a == b
    assert nvar == st.n_maps + 2
    glg = open(tmp_path / "out.glg").read()
    assert "Successful Completion" in glg[-200:]


def test_write_maps_without_binary_flag(tmp_path):
    job, gpf = _job(tmp_path, "small")
    st = job.state()
    maps = np.zeros((st.n_maps, st.points), np.float32)
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        job.write_maps(maps)
    finally:
        os.chdir(cwd)
    job.close()
    assert not list(tmp_path.glob("*.bin")) and not (tmp_path / "small.maps.bin.fld").exists()
    for fn in cases.map_names(os.path.join(tmp_path, gpf)):
        lines = open(tmp_path / fn).read().split("\n")
        assert lines[:6] == GOLDEN["small"]["maps"][fn]["header"]
        assert set(lines[6:-1]) == {"0"}
