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


def test_formatter_fast_path_edges_match_printf():
    """Below |n| = 16,384,000 (n = floor(1000 e + 0.5)) the formatter prints
    n / 1000 directly, without the float conversion.  Values at and around
    that limit, at the half-thousandths where rounding turns, and their
    float neighbours, against printf."""
    base = [16383.999, 16384.0, 16384.001, 16383.9995, 16384.0005, 8191.9995, 8192.0005, 4095.9995, 1e-4, 5e-4,
            1.0005, 2.0625, 1023.4995, 12345.6785, 99999.9995]
    k = np.arange(-200, 200)
    grid = np.concatenate([(np.arange(16380000, 16388000, 7) + 0.5) / 1000.0, (k + 0.5) / 1000.0,
                           (np.arange(0, 16384000, 4093) + 0.5) / 1000.0])
    v = np.concatenate([np.array(base), grid]).astype(np.float32)
    v = np.concatenate([v, -v])
    # the float neighbours of each value
    bits = v.view(np.uint32).astype(np.int64)
    near = np.concatenate([(bits + d).astype(np.uint32).view(np.float32) for d in (-2, -1, 1, 2)])
    v = np.concatenate([v, near[np.isfinite(near)]])
    mine = ag.format_map_values(v)
    ref = ol.format_map(v.astype(np.float64))
    assert mine == ref
