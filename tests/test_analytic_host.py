"""Analytic (differentiable) mode: host-side checks (no GPU).

* The FP64 analytic model of the oracle is pinned to the reference's table
  path (orc_maps, itself pinned to the reference binary): away from the
  pieces the analytic model changes on purpose -- the 1/128 A bins, the
  discrete smoothing window and the 7.5-8 A switch in place of the hard
  cutoff -- the two agree closely, and the e map, which has no cutoff or
  window, agrees to the bin quantisation of eps(r).
* The gradient through the bond vectors (ag_bond_vectors_vjp) against FP64
  central differences of ag_bond_vectors.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

import autogrid_amd as ag
import cases
import oracle_lib as ol


def _state(name, tmp_path):
    gpf = cases.materialize(name, str(tmp_path))
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        return ag.Job(gpf).prepare().state()
    finally:
        os.chdir(cwd)


def test_analytic_oracle_matches_table_path(tmp_path):
    st = _state("small", tmp_path)
    rng = np.random.default_rng(7)
    pts = np.sort(rng.choice(st.points, 600, replace=False))
    an = ol.analytic_maps(st, pts)
    z = np.unique(pts // (st.n1[0] * st.n1[1]))
    tab = np.zeros_like(an)
    full = ol.maps(st, int(z.min()), int(z.max()) + 1).reshape(st.n_maps, -1)
    tab[:] = full[:, pts]
    e = st.elec_index
    # e map: eps(r) at the exact distance instead of at floor(128 r)/128.
    # Each term can move by at most |q| c 332 / max(r, .5) |d(1/eps)/dr| / 128
    # (the derivative bounded by its larger end value over the bin).
    ix, iy, iz = pts % st.n1[0], (pts // st.n1[0]) % st.n1[1], pts // (st.n1[0] * st.n1[1])
    c = np.stack([ix - st.ne[0], iy - st.ne[1], iz - st.ne[2]], axis=1) * st.spacing
    r = np.linalg.norm(st.coord[None, :, :] - c[:, None, :], axis=2)
    rb = np.floor(r * 128.0) / 128.0

    def inv_eps(x):
        lam, e0, A, k = 0.003627, 78.4, -8.5525, 7.7839
        B = e0 - A
        return 1.0 / (A + B / (1.0 + k * np.exp(-lam * B * x)))

    h = 1e-6
    deriv = np.maximum(np.abs(inv_eps(r + h) - inv_eps(r - h)), np.abs(inv_eps(rb + h) - inv_eps(rb - h))) / (2 * h)
    bound = (np.abs(st.charge)[None, :] * st.coeff_estat * 332.0 / np.maximum(r, 0.5) * deriv * (r - rb)).sum(axis=1)
    diff = np.abs(an[e] - tab[e])
    assert np.all(diff <= 1.01 * bound + 1e-9), float(np.max(diff - bound))
    assert np.max(diff) > 0.0
    # affinity / d maps: the switch and the continuous window move values a
    # little; bulk agreement stays tight where the maps are not repulsive
    for m in range(st.n_maps):
        if m == e:
            continue
        ok = np.abs(tab[m]) < 5.0
        d = np.abs(an[m][ok] - tab[m][ok])
        assert np.median(d) < 0.02 and np.percentile(d, 95) < 0.15, (m, np.median(d), np.percentile(d, 95))
        assert np.corrcoef(an[m][ok], tab[m][ok])[0, 1] > 0.99


def test_bond_vector_vjp_against_finite_differences(tmp_path):
    st = _state("small", tmp_path)
    rng = np.random.default_rng(11)
    n = st.n_atoms
    g_rv = rng.normal(size=(n, 3))
    g_rv2 = rng.normal(size=(n, 3))
    got = ag.bond_vectors_vjp(st, st.coord, g_rv, g_rv2)

    def f(coord):
        _, rv, rv2, _ = ag.bond_vectors(st, coord, st.type, st.hbond)
        return float(np.sum(g_rv * rv) + np.sum(g_rv2 * rv2))

    # every atom that defines some bond vector, a few components each
    idx = np.nonzero(np.abs(got).sum(axis=1) > 0)[0]
    assert idx.size > 50
    h = 1e-6
    for i in rng.choice(idx, 40, replace=False):
        for k in range(3):
            c1 = st.coord.copy()
            c1[i, k] += h
            c2 = st.coord.copy()
            c2[i, k] -= h
            fd = (f(c1) - f(c2)) / (2 * h)
            assert abs(fd - got[i, k]) <= 1e-6 * max(1.0, abs(fd)), (i, k, fd, got[i, k])


def test_analytic_weights_from_library(tmp_path):
    st = _state("small", tmp_path)
    # AD4.1_bound.dat FE_coeff_* lines
    assert st.weights == pytest.approx((0.1662, 0.1209, 0.1406, 0.1322))
