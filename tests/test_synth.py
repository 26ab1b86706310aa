"""Seeded workload generator: determinism and the stated geometry."""
from __future__ import annotations

import numpy as np

from autogrid_amd import synth


def test_receptor_is_deterministic_and_well_formed():
    a = synth.make_receptor(600, 20.0, 11, center=(1.0, 2.0, 3.0))
    b = synth.make_receptor(600, 20.0, 11, center=(1.0, 2.0, 3.0))
    assert np.array_equal(a[0], b[0]) and list(a[1]) == list(b[1]) and np.array_equal(a[2], b[2])
    xyz, types, q = a
    assert len(xyz) == 600 and abs(q.sum()) < 1e-9
    assert np.all(np.linalg.norm(xyz - np.array([1.0, 2.0, 3.0]), axis=1) <= 20.0 + 1.0 + 1e-9)
    types = np.array(types)
    assert set(types) <= {"C", "N", "OA", "SA", "HD"}
    for i in np.nonzero(types == "HD")[0]:
        assert types[i - 1] in ("N", "OA")
        assert abs(np.linalg.norm(xyz[i] - xyz[i - 1]) - 1.0) < 1e-9
        d = np.linalg.norm(xyz - xyz[i], axis=1)
        d[[i, i - 1]] = 99
        assert d.min() >= 1.5 - 1e-9
    heavy = np.nonzero(types != "HD")[0]
    hx = xyz[heavy]
    dd = np.linalg.norm(hx[:, None] - hx[None], axis=-1)
    idx = np.arange(len(heavy))
    far = np.abs(idx[:, None] - idx[None]) >= 3
    assert dd[far].min() >= 2.9 - 1e-9
