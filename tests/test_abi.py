"""The C-ABI library loads and exports every entry point of include/*.h."""
from __future__ import annotations

import ctypes as C
import glob
import os
import re

import autogrid_amd as ag
from autogrid_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(ag_[a-z_0-9]+)\s*\(", text, re.M):
            names.add(m.group(1))
    return names


def test_every_declared_symbol_is_exported():
    lib = _abi.lib()
    names = declared_functions()
    assert len(names) >= 15, names
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_abi_version():
    assert _abi.lib().ag_abi_version() == 3


def test_context_on_a_host_without_gpu_fails_cleanly():
    h = C.c_void_p()
    st = _abi.lib().ag_context_create(C.byref(h), 0)
    if st == 0:       # a device is present: nothing to check here
        _abi.lib().ag_context_destroy(h)
        return
    assert st == 4    # AG_ERR_DEVICE, never a crash or exit
    assert b"no HIP device" in _abi.lib().ag_last_error() or b"HIP" in _abi.lib().ag_last_error()


def test_null_arguments_are_rejected():
    L = _abi.lib()
    assert L.ag_maps_plain(None, None, None, None, None, None, None, None, 0) == 1
    assert L.ag_plan_execute(None, None, None, None) == 1
    assert L.ag_job_open(None, None, None, 0, None) == 1
