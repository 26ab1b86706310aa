"""The allocation-free entry point (ag_maps_plain_noalloc) on the GPU.

Its map set must equal ag_maps_plain's with device outputs bit for bit: the
same kernels run on the same inputs, only the device buffers come from the
library's fixed scratch array instead of allocations.  Two calls in a row,
the second on a smaller receptor and grid, check that the array's leftover
contents from the first call do not leak into the second; a third call, on a
map set whose buffers exceed the array, must be refused with
AG_ERR_UNSUPPORTED and leave its outputs untouched.
"""
from __future__ import annotations

import ctypes as C
import dataclasses

import numpy as np
import pytest

import autogrid_amd as ag
import parity
from autogrid_amd import _abi

pytestmark = pytest.mark.gpu


def _hip():
    return C.CDLL("libamdhip64.so.7")


class Dev:
    def __init__(self, nbytes, fill=0xA5):
        self.hip = _hip()
        self.p = C.c_void_p()
        self.n = nbytes
        assert self.hip.hipMalloc(C.byref(self.p), C.c_size_t(nbytes)) == 0
        assert self.hip.hipMemset(self.p, fill, C.c_size_t(nbytes)) == 0
        assert self.hip.hipDeviceSynchronize() == 0

    def get(self, dtype=np.float32):
        out = np.empty(self.n // np.dtype(dtype).itemsize, dtype)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p, C.c_size_t(self.n), 2) == 0
        return out

    def free(self):
        self.hip.hipFree(self.p)


def _both(ctx, st):
    """(library path, allocation-free path) map sets of st, each into its
    own device buffers filled with a byte pattern beforehand; fmap / cmap too
    when the map set asks for them."""
    nb = 4 * st.n_maps * st.points
    runs = []
    for noalloc in (False, True):
        m = Dev(nb)
        f = Dev(4 * st.points) if st.floating_grid else None
        c = Dev(4 * st.points) if st.constriction_grid else None
        try:
            if noalloc:
                ctx.maps_plain_noalloc(st, m.p.value, f.p.value if f else None, c.p.value if c else None)
            else:
                rec, grid, ms, tab, _ = st.structs()
                ag._abi.check(ag._abi.lib().ag_maps_plain(ctx.handle, C.byref(rec), C.byref(grid), C.byref(ms),
                                                          C.byref(tab), m.p, f.p if f else None,
                                                          c.p if c else None, 1))
            runs.append((m.get(), f.get() if f else None, c.get(np.int32) if c else None))
        finally:
            for b in (m, f, c):
                if b:
                    b.free()
    return runs


def test_noalloc_equals_library_path_twice(tmp_path):
    big, _ = parity.load_case("all13", str(tmp_path / "a"))
    big = dataclasses.replace(big, floating_grid=1, constriction_grid=1, constriction_cutoff=6.5, _keep=[])
    small, _ = parity.load_case("hsg1_sm", str(tmp_path / "b"))
    assert small.n_atoms < big.n_atoms and small.points < big.points
    with ag.Context(0) as ctx:
        for name, st in (("all13 (fmap, cmap)", big), ("hsg1_sm", small)):
            print(f"[gpu] noalloc vs ag_maps_plain: {name}, {st.n_atoms} atoms, {st.points} points", flush=True)
            (m0, f0, c0), (m1, f1, c1) = _both(ctx, st)
            assert np.array_equal(m0.view(np.uint32), m1.view(np.uint32)), name
            if f0 is not None:
                assert np.array_equal(f0.view(np.uint32), f1.view(np.uint32)), name
                assert np.array_equal(c0, c1), name


def test_noalloc_refuses_a_map_set_beyond_its_scratch(tmp_path):
    """A tree-build map set with collisions keeps one byte per point: on a
    329^3 grid that alone passes the 32 MiB scratch, so the call is refused
    before any kernel runs, and the output keeps its fill pattern."""
    st, _ = parity.load_case("hsg1_sm", str(tmp_path))
    n = 329
    st = dataclasses.replace(st, n1=(n, n, n), ne=(n // 2,) * 3, bhtree=1, _keep=[])
    assert st.points > _abi.AG_NOALLOC_SCRATCH_BYTES
    print(f"[gpu] noalloc refusal: {st.points} points, tree build", flush=True)
    with ag.Context(0) as ctx:
        m = Dev(4 * st.n_maps * st.points)
        try:
            with pytest.raises(ag.AgError) as e:
                ctx.maps_plain_noalloc(st, m.p.value)
            assert e.value.status == _abi.AG_ERR_UNSUPPORTED and "MiB" in str(e.value)
            back = m.get().view(np.uint8)
            assert np.all(back == 0xA5)
        finally:
            m.free()
