"""Map value text: the engine's formatter against printf (via the oracle)."""
from __future__ import annotations

import numpy as np

import autogrid_amd as ag
import oracle_lib as ol


def _cases():
    rng = np.random.default_rng(7)
    v = [rng.normal(0, s, 4000) for s in (1e-4, 1e-3, 0.1, 1, 30, 3000, 1e5)]
    ties = (np.arange(-5000, 5000) + 0.5) / 1000.0
    special = np.array([0.0, -0.0, 1e-4, -1e-4, 9.99e-5, 0.0004999, -0.0005, 0.0005, 1e5, -1e5, 2e5, 99999.9995,
                        8191.9995, 16777216.0, 3.0e10, -7.25e8])
    return np.concatenate(v + [ties, special]).astype(np.float32)


def test_formatter_matches_printf():
    v = _cases()
    mine = ag.format_map_values(v)
    ref = ol.format_map(v.astype(np.float64))
    assert mine == ref
