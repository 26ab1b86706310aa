"""Analytic (differentiable) mode: host-side checks (no GPU).

* The FP64 analytic model of the oracle, one pair at a time: on a one-atom
  receptor over r = 0.1-8.3 A, E(r) equals the brute-force evaluation of
  the min-window smoothing's definition on the textbook 12-6 / 12-10
  potential (capped at 1e5), which pins the closed form of the window on
  the repulsive wall, in the well and beyond it; it lies within the
  Lipschitz bound of the reference's table at bin floor(128 r); the 7.5-8 A
  switch is the C2 quintic; the d map's Gaussian matches the reference's
  sol_fn to its bin bound.
* The whole analytic map set against the reference's table path (orc_maps,
  itself pinned to the reference binary) at every sampled point, within a
  per-point bound summed from each atom's Lipschitz constants: the two
  share the per-point assembly and differ only in E and sol at the exact
  distance and in the switch; the bound is reached to within 3% on every
  map.  The e map, which has no cutoff or window, agrees to the bin
  quantisation of eps(r).
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


def _phi(pair, x):
    """The AD4 pair potential of a (12, 6) or (12, 10) pair in its textbook
    form, minimum -eps at Rij: eps((R/x)^12 - 2(R/x)^6) and
    eps(5(R/x)^12 - 6(R/x)^10); and its derivative."""
    R, eps, a, b = float(pair["Rij"]), float(pair["eps"]), int(pair["xA"]), int(pair["xB"])
    rho = R / x
    if (a, b) == (12, 6):
        return eps * (rho ** 12 - 2 * rho ** 6), eps * (-12 * rho ** 12 + 12 * rho ** 6) / x
    assert (a, b) == (12, 10)
    return eps * (5 * rho ** 12 - 6 * rho ** 10), eps * (-60 * rho ** 12 + 60 * rho ** 10) / x


def _lipschitz_table(st):
    """L[m, t, k]: a Lipschitz constant of the pair energy near both window
    minima of bin k.  The table's discrete min-window (main.cpp:2200-2216,
    W = (int)(64 r_smooth) bins) takes its minimum at some bin d* of
    [k - W, k + W]; the continuous window [r - w, r + w] of an r in bin k
    at some x*.  Each window has a point within 1/128 A of the other's
    argmin, so the two minima differ by at most L / 128 with L = max |Phi'|
    over [min(x*, d*) - 1/128, max(x*, d*) + 1/128].  d* is found over the
    bins, x* by brute force over the window at both ends of the bin (the
    argmin of a single-well potential moves monotonically with r)."""
    W = int(st.r_smooth * 128 / 2)
    w = W / 128.0
    M, T = st.pairs.shape
    k = np.arange(1024)
    sub = np.linspace(-1.0, 1.0, 1025)
    L = np.zeros((M, T, 1024))
    jj = np.clip(k[:, None] + np.arange(-W, W + 1)[None, :], 0, 1023)
    for m in range(M):
        for t in range(T):
            if st.maps[m]["is_covalent"]:
                continue
            pair = st.pairs[m, t]
            xj = np.maximum(jj / 128.0, 1e-3)
            dstar = xj[k, np.argmin(np.minimum(_phi(pair, xj)[0], 1e5), axis=1)]
            ends = []
            for r0 in (k / 128.0, (k + 1) / 128.0):
                xs = np.maximum(r0[:, None] + w * sub[None, :], 1e-3)
                ends.append(xs[k, np.argmin(_phi(pair, xs)[0], axis=1)])
            lo = np.maximum(np.minimum.reduce([dstar, ends[0], ends[1]]) - 1 / 128.0 - 2e-3, 1e-3)
            hi = np.maximum.reduce([dstar, ends[0], ends[1]]) + 1 / 128.0 + 2e-3
            x = lo[:, None] + (hi - lo)[:, None] * np.linspace(0.0, 1.0, 33)[None, :]
            L[m, t] = np.max(np.abs(_phi(pair, x)[1]), axis=1) * 1.02
    return L


def _one_atom_line(st, rtype, offset, n=1025, h=0.002):
    """st reduced to one atom of receptor type rtype (no charge, volume or
    solvation parameter, no H-bond class) and a line of n points along x,
    spacing h, whose distances to the atom run from about offset to
    offset + (n - 1) h; every map made a non-H-bonding, desolvation-free
    map, so each map value is the atom's pair energy E(r) alone."""
    import dataclasses
    maps = st.maps.copy()
    maps["is_hbonder"] = 0
    maps["hbond"] = 0
    maps["solpar_probe"] = 0.0
    maps["vol_probe"] = 0.0
    one = dataclasses.replace(
        st, coord=np.array([[-offset, 0.00123, 0.00071]]), charge=np.zeros(1), vol=np.zeros(1), solpar=np.zeros(1),
        type=np.array([rtype], np.int32), hbond=np.zeros(1, np.int32), rexp=np.zeros(1, np.int32),
        rvector=np.zeros((1, 3)), rvector2=np.zeros((1, 3)), disorder=np.zeros(1, np.uint8), n1=(n, 1, 1),
        ne=(0, 0, 0), spacing=h, maps=maps, _keep=[])
    r = np.sqrt((np.arange(n) * h + offset) ** 2 + 0.00123 ** 2 + 0.00071 ** 2)
    return one, r


def test_analytic_pair_energy_is_the_smoothing_definition(tmp_path):
    """One atom, every (map, receptor type) pair, r from 0.1 to 8.3 A in
    2 mA steps.  The analytic E(r) is checked against

    * the definition of the min-window smoothing evaluated by brute force:
      min over [r - w, r + w] (2,001 samples, plus Rij when inside) of the
      textbook pair potential, capped at EINTCLAMP = 1e5 -- which pins the
      closed form s(r) = r + w / r* / r - w, its left branch on the
      repulsive wall included;
    * the reference's table at bin floor(128 r) (tables pinned to the
      reference binary), within L / 128, L the largest |dPhi/dr| over the
      windows (Lipschitz bound of the discretisation);
    * the switch on 7.5-8 A: E_analytic / E_brute = 1 - (10u^3 - 15u^4 +
      6u^5), u = (r - 7.5) / 0.5 (the C2 quintic, fixed by its six end
      conditions), with value and slope continuous at both ends, and
      exactly 0 from 8 A on."""
    st = _state("small", tmp_path)
    M, T = st.pairs.shape
    w = np.floor(st.r_smooth * 64) / 128
    L = _lipschitz_table(st)
    tabs = ol.tables(st)[0]
    for t in range(T):
        an_all, r_all = [], []
        for off in (0.1, 2.148, 4.196, 6.244):
            one, r = _one_atom_line(st, t, off)
            an_all.append(ol.analytic_maps(one, np.arange(one.points)))
            r_all.append(r)
        an = np.concatenate(an_all, axis=1)
        r = np.concatenate(r_all)
        for m in range(M):
            if st.maps[m]["is_covalent"]:
                continue
            pair = st.pairs[m, t]
            xs = np.maximum(r[:, None] + w * np.linspace(-1.0, 1.0, 2001)[None, :], 1e-3)
            bf = np.min(_phi(pair, xs)[0], axis=1)
            inside = np.abs(r - pair["Rij"]) <= w
            bf[inside] = np.minimum(bf[inside], -float(pair["eps"]))
            bf = np.minimum(bf, 1.0e5)
            got = an[m]
            core = r <= 7.5
            assert np.allclose(got[core], bf[core], rtol=1e-9, atol=1e-12), (m, t)
            # the reference table, within the discretisation bound
            k = np.minimum((r * 128).astype(int), 1023)
            tab = tabs[t, m, k].astype(np.float64)
            ok = r <= 8.0
            assert np.all(np.abs(bf[ok] - tab[ok]) <= L[m, t, k[ok]] / 128 + 1e-6 * np.abs(tab[ok]) + 1e-9), (m, t)
            # switch
            sw = (r > 7.5) & (r < 8.0) & (np.abs(bf) > 1e-12)
            u = (r[sw] - 7.5) / 0.5
            assert np.allclose(got[sw] / bf[sw], 1 - (10 * u ** 3 - 15 * u ** 4 + 6 * u ** 5), rtol=1e-9, atol=1e-12)
            assert np.all(got[r >= 8.0] == 0.0)
    # the quintic's end conditions (value, slope, curvature) -- what fixes it
    S = lambda x: 1 - (10 * x ** 3 - 15 * x ** 4 + 6 * x ** 5)  # noqa: E731
    e = 1e-4
    assert S(0.0) == 1.0 and S(1.0) == 0.0
    assert abs(S(e) - 1) < 1e-10 and abs(S(1 - e)) < 1e-10


def test_analytic_desolvation_one_atom(tmp_path):
    """One atom with volume 20 A^3: the d map is solpar_q vol sol(r) S(r) with
    sol the Gaussian of main.cpp:2272-2276 at the exact r; against the
    reference's sol_fn at bin floor(128 r) within |d sol/dr| / 128."""
    import dataclasses
    st = _state("small", tmp_path)
    sig2 = 3.6 ** 2
    d = st.dsolv_index
    for off in (0.1, 2.148, 4.196, 6.244):
        one, r = _one_atom_line(st, 0, off)
        one = dataclasses.replace(one, vol=np.array([20.0]), _keep=[])
        an = ol.analytic_maps(one, np.arange(one.points))[d]
        k = (r * 128).astype(int)
        tab = st.solpar_q * 20.0 * st.sol_fn[k]
        slope = st.solpar_q * 20.0 * st.coeff_desolv * (k + 1) / 128 / sig2 * np.exp(-(k / 128) ** 2 / (2 * sig2))
        core = r <= 7.5
        assert np.all(np.abs(an[core] - tab[core]) <= slope[core] / 128 * 1.01 + 1e-15)
        exact = st.solpar_q * 20.0 * st.coeff_desolv * np.exp(-r ** 2 / (2 * sig2))
        assert np.allclose(an[core], exact[core], rtol=1e-12)
        assert np.all(an[r >= 8.0] == 0.0)


def _case_factor(mh, is_hb, ah):
    """Lipschitz constant of a map's per-atom term in the pair energy E
    (main.cpp:3286-3374): plain add 1; E Hramp (racc + (1 - racc) rsph(E))
    at most 2 (rsph = clamp(E / 100, 0, 1)); the min/max pair hmin +
    max(hmax, 0) of such terms at most 2 x 2."""
    don = ah in (1, 2, 6)
    if not is_hb:
        return 1.0
    if mh in (3, 5, 6) and don:
        return 2.0
    if (mh in (4, 6) and don) or (mh in (1, 2, 6) and (ah > 2 or ah == 6)):
        return 4.0
    return 1.0


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
    # affinity and d maps: a per-point bound from the Lipschitz constants of
    # each atom's term.  Both models share the per-point assembly
    # (main.cpp:2985-3431) and differ only in E and sol: E at bin floor(128 r)
    # with the discrete window against E at r with the continuous one
    # (|dE| <= L / 128), sol likewise (|d sol| <= max |sol'| / 128), and,
    # for 7.5 < r <= 8, the switch, which moves a term by at most its size.
    L = _lipschitz_table(st)
    tabs = ol.tables(st)[0]
    sig2 = 3.6 ** 2
    within = r <= 8.0
    k = np.minimum((r * 128).astype(int), 8191)
    kt = np.minimum(k, 1023)
    shell = (r > 7.5) & within
    gk = np.exp(-(k / 128.0) ** 2 / (2 * sig2))
    dsol = st.coeff_desolv * ((k + 1) / 128.0 / sig2 * gk / 128.0 + shell * gk)
    assert not st.separate_desolvation
    d = st.dsolv_index
    checked = 0
    for m in range(st.n_atom_maps):
        if st.maps[m]["is_covalent"]:
            continue
        mp = st.maps[m]
        c = np.array([_case_factor(int(mp["hbond"]), int(mp["is_hbonder"]), int(h)) for h in st.hbond])
        Lm = L[m][st.type[None, :], kt]
        Et = tabs[st.type[None, :], m, kt].astype(np.float64)
        dE = Lm / 128.0 + shell * (np.abs(Et) + Lm / 128.0)
        coef = np.abs(mp["solpar_probe"] * st.vol + (st.solpar + st.solpar_q * np.abs(st.charge)) * mp["vol_probe"])
        bound = np.sum(within * (c[None, :] * dE + coef[None, :] * dsol), axis=1)
        diff = np.abs(an[m] - tab[m])
        assert np.all(diff <= bound * (1 + 1e-9) + 1e-9), (m, float(np.max(diff - bound)))
        # not vacuous: away from atoms the bound is tight
        assert np.median(bound) < 0.1 and np.max(diff / np.maximum(bound, 1e-30)) > 0.5, (m, float(np.median(bound)))
        checked += 1
    bound_d = np.sum(within * (st.solpar_q * st.vol)[None, :] * dsol, axis=1)
    assert np.all(np.abs(an[d] - tab[d]) <= bound_d * (1 + 1e-9) + 1e-12)
    assert checked == st.n_atom_maps


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
