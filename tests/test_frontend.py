"""Front end: GPF semantics, host tables and bond vectors, CLI error contract.

The engine's host precompute (tables, bond vectors, disorder) must equal the
oracle's bit for bit; the oracle equals the reference (test_oracle.py).
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

GOLDEN = json.load(open(os.path.join(cases.GOLDEN, "golden.json")))
SUCCESS = [n for n in cases.ALL_CASES if GOLDEN[n]["success"]]
FAILURE = [n for n in cases.ALL_CASES if not GOLDEN[n]["success"]]


@pytest.mark.parametrize("name", SUCCESS)
def test_host_precompute_equals_oracle(name, tmp_path):
    st, _ = parity.load_case(name, str(tmp_path))
    ost = ol.oracle_state(st)
    for k in ("e_vdw_hb", "sol_fn", "r_eps", "rexp", "rvector", "rvector2", "disorder"):
        a, b = getattr(st, k), getattr(ost, k)
        assert a.shape == b.shape and np.array_equal(a, b), k
    assert st.n_maps == len(GOLDEN[name]["maps"])


@pytest.mark.parametrize("name", FAILURE)
def test_bad_gpf_is_rejected(name, tmp_path):
    gpf = cases.materialize(name, str(tmp_path))
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        with pytest.raises(ag.AgError) as e:
            ag.Job(gpf).prepare()
        assert e.value.status == 1
    finally:
        os.chdir(cwd)


@pytest.mark.parametrize("name", FAILURE)
def test_cli_failure_contract(name, tmp_path):
    """Exit 1 and 'Unsuccessful Completion' at the end of the log (main.cpp:173-176)."""
    gpf = cases.materialize(name, str(tmp_path))
    p = subprocess.run([ag.CLI_PATH, "-p", gpf, "-l", "out.glg"], cwd=tmp_path, capture_output=True, timeout=60)
    assert p.returncode == 1
    tail = "".join(open(tmp_path / "out.glg").read().splitlines(True)[-10:])
    assert "Unsuccessful Completion" in tail and "ERROR" in tail
    assert b"ERROR" in p.stderr


@pytest.mark.parametrize("name", SUCCESS)
def test_cli_writes_reference_headers(name, tmp_path):
    """The .xyz text, the .fld, the six map-header lines and the log's
    'Calculating N grids over P elements' line are written before the
    device's outcome is looked at, and must equal the reference's bytes
    (main.cpp:1160-1164, 1897-1956, 2740-2742, 2752-2776)."""
    gpf = cases.materialize(name, str(tmp_path))
    subprocess.run([ag.CLI_PATH, "-p", gpf, "-l", "out.glg"], cwd=tmp_path, capture_output=True, timeout=60)
    g = GOLDEN[name]
    for m in cases.map_names(os.path.join(tmp_path, gpf)):
        assert open(tmp_path / m).read().split("\n")[:6] == g["maps"][m]["header"]
    if g.get("fld"):
        fld = [ln for ln in open(tmp_path / f"{name}.maps.fld").read().split("\n") if not ln.startswith("# Created by")]
        assert fld == g["fld"]
    assert g["xyz"] is not None
    xyz = cases.xyz_name(os.path.join(tmp_path, gpf))
    assert open(tmp_path / xyz).read() == g["xyz"]
    assert cases.glg_record(open(tmp_path / "out.glg").read())["calculating"] == g["glg"]["calculating"]


def test_gpf_keywords_are_case_insensitive_and_comments_ignored(tmp_path):
    gpf = cases.materialize("hsg1_sm", str(tmp_path))
    text = open(tmp_path / gpf).read()
    mangled = "# a comment line\n\n" + "\n".join(
        ln.split()[0].upper() + ln[len(ln.split()[0]):] if ln.split() else ln for ln in text.splitlines()) + "\n"
    (tmp_path / "upper.gpf").write_text(mangled)
    a, _ = parity.load_case("hsg1_sm", str(tmp_path / "a"))
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        b = ag.Job("upper.gpf").prepare().state()
    finally:
        os.chdir(cwd)
    assert np.array_equal(a.e_vdw_hb, b.e_vdw_hb) and np.array_equal(a.coord, b.coord)


def test_unknown_receptor_type_is_fatal(tmp_path):
    gpf = cases.materialize("hsg1_sm", str(tmp_path))
    pdb = (tmp_path / "hsg1_sm.pdbqt").read_text().splitlines(True)
    i = next(k for k, ln in enumerate(pdb) if ln.startswith("ATOM"))
    pdb[i] = pdb[i][:77] + "Xx\n"
    (tmp_path / "hsg1_sm.pdbqt").write_text("".join(pdb))
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        with pytest.raises(ag.AgError, match="unknown type"):
            ag.Job(gpf)
    finally:
        os.chdir(cwd)


def test_npts_rounds_down_to_even(tmp_path):
    gpf = cases.materialize("hsg1_sm", str(tmp_path))
    text = open(tmp_path / gpf).read().replace("npts 4 4 4", "npts 5 7 3")
    (tmp_path / gpf).write_text(text)
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        st = ag.Job(gpf).state()
    finally:
        os.chdir(cwd)
    assert st.n1 == (5, 7, 3) and st.ne == (2, 3, 1)


def test_cli_reports_the_device_failure_after_valid_inputs(tmp_path):
    """The drop-in opens the device on a thread while it reads the inputs.
    On a host without a GPU and with valid inputs, the device's failure is
    what ends the run: exit 1, the HIP error and 'Unsuccessful Completion'
    in the log, after the headers of every map were written."""
    import ctypes as C
    from autogrid_amd import _abi
    h = C.c_void_p()
    if _abi.lib().ag_context_create(C.byref(h), 0) == 0:
        _abi.lib().ag_context_destroy(h)
        pytest.skip("a GPU is present")
    gpf = cases.materialize("small", str(tmp_path))
    p = subprocess.run([ag.CLI_PATH, "-p", gpf, "-l", "out.glg"], cwd=tmp_path, capture_output=True, timeout=60)
    assert p.returncode == 1
    log = open(tmp_path / "out.glg").read()
    tail = "".join(log.splitlines(True)[-10:])
    assert "Unsuccessful Completion" in tail and "HIP" in tail
    for m in cases.map_names(os.path.join(tmp_path, gpf)):
        assert len(open(tmp_path / m).read().split("\n")) >= 6
