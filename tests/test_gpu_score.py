"""Trilinear ligand scoring kernel (ag_scorer_*) on the GPU.

Checks, on seeded maps and poses:
* scores against an FP64 host evaluation of the same trilinear sums
  (poses.splat), relative 1e-5 (FP32 interpolation, FP64 pose sums);
* exactness on linear fields: trilinear interpolation reproduces
  a x + b y + c z + d, so every atom's energy and gradient follow in closed
  form;
* per-atom gradients against FP64 central differences of the host
  evaluation;
* the adjoint against the host cotangent S^T g, the dot-product identity
  <S^T g, m> = <g, S m>, and bitwise repeatability;
* atoms whose cell leaves the grid score the penalty with zero gradient.
"""
from __future__ import annotations

import ctypes as C
import types

import numpy as np
import pytest

import autogrid_amd as ag
from autogrid_amd import poses as ps

pytestmark = pytest.mark.gpu


class Dev:
    def __init__(self, nbytes):
        self.hip = C.CDLL("libamdhip64.so.7")
        self.p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.p), C.c_size_t(max(nbytes, 16))) == 0
        self.n = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a, np.float32)
        assert a.nbytes <= self.n
        assert self.hip.hipMemcpy(self.p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0

    def get(self, n):
        out = np.empty(n, np.float32)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p, C.c_size_t(out.nbytes), 2) == 0
        return out

    def free(self):
        self.hip.hipFree(self.p)


def _grid(n1=(23, 19, 17), ne=(11, 9, 8), spacing=0.375, n_maps=5, e=3, d=4):
    return types.SimpleNamespace(n1=list(n1), ne=list(ne), spacing=spacing, n_maps=n_maps, elec_index=e,
                                 dsolv_index=d, points=int(np.prod(n1)))


def _poses(st, n_poses=12, seed=7, margin=0.05):
    rng = np.random.default_rng(seed)
    lo = np.array([-st.ne[k] * st.spacing for k in range(3)])
    hi = np.array([(st.n1[k] - 1 - st.ne[k]) * st.spacing for k in range(3)])
    out = []
    for _ in range(n_poses):
        n = int(rng.integers(5, 40))
        x = rng.uniform(lo + margin, hi - margin, size=(n, 3))
        mi = rng.integers(-1, 3, size=n)          # -1: no affinity map
        q = rng.normal(0, 0.3, size=n)
        out.append((mi, q, x))
    return out


def _host_scores(st, poses, maps):
    r, c, v = ps.splat(st, poses)
    s = np.zeros(len(poses))
    np.add.at(s, r, v * maps.reshape(-1)[c])
    return s


def _run(st, poses, maps, with_grad=True):
    n_at = sum(len(p[0]) for p in poses)
    dm, ds, dg = Dev(4 * maps.size), Dev(4 * len(poses)), Dev(12 * n_at)
    try:
        dm.put(maps)
        with ag.Context(0) as ctx, ag.Scorer(ctx, st, poses, outside=123.5) as sc:
            sc.score(dm.p.value, ds.p.value, dg.p.value if with_grad else None)
            return ds.get(len(poses)).astype(np.float64), dg.get(3 * n_at).reshape(-1, 3).astype(np.float64)
    finally:
        dm.free(); ds.free(); dg.free()


def test_scores_against_host():
    print("[gpu] scorer vs host", flush=True)
    st = _grid()
    rng = np.random.default_rng(1)
    maps = rng.normal(0, 2.0, size=(st.n_maps, st.points)).astype(np.float32)
    poses = _poses(st)
    got, _ = _run(st, poses, maps)
    ref = _host_scores(st, poses, maps.astype(np.float64))
    assert np.allclose(got, ref, rtol=1e-5, atol=1e-4), np.max(np.abs(got - ref))


def test_linear_fields_exact():
    print("[gpu] scorer linear fields", flush=True)
    st = _grid()
    n1x, n1y, n1z = st.n1
    k, j, i = np.meshgrid(np.arange(n1z), np.arange(n1y), np.arange(n1x), indexing="ij")
    x = (i - st.ne[0]) * st.spacing
    y = (j - st.ne[1]) * st.spacing
    z = (k - st.ne[2]) * st.spacing
    coef = np.array([[0.5, -1.0, 0.25, 2.0], [1.5, 0.5, -0.75, -1.0], [-0.25, 0.125, 1.0, 0.5],
                     [0.75, -0.5, 0.5, 1.0], [-1.0, 0.25, -0.5, 0.25]])
    maps = np.stack([(a * x + b * y + c * z + d).reshape(-1) for a, b, c, d in coef]).astype(np.float32)
    poses = _poses(st, seed=11)
    got, grad = _run(st, poses, maps)
    o = 0
    for p, (mi, q, xyz) in enumerate(poses):
        want = 0.0
        for a in range(len(mi)):
            w = np.zeros(st.n_maps)
            if mi[a] >= 0:
                w[mi[a]] += 1.0
            w[st.elec_index] += q[a]
            w[st.dsolv_index] += abs(q[a])
            lin = w @ coef                       # this atom's field a x + b y + c z + d
            want += lin[:3] @ xyz[a] + lin[3]
            assert np.allclose(grad[o + a], lin[:3], rtol=1e-4, atol=1e-4), (p, a, grad[o + a], lin[:3])
        o += len(mi)
        assert abs(got[p] - want) <= 1e-5 * max(1.0, abs(want)) * len(mi), (p, got[p], want)


def test_gradients_against_fd():
    print("[gpu] scorer gradient vs FD", flush=True)
    st = _grid()
    rng = np.random.default_rng(2)
    maps = rng.normal(0, 2.0, size=(st.n_maps, st.points))
    poses = _poses(st, n_poses=3, seed=5)
    _, grad = _run(st, poses, maps.astype(np.float32))
    h = 1e-6
    o = 0
    for p, (mi, q, xyz) in enumerate(poses):
        for a in range(len(mi)):
            g = (xyz[a] + st.ne * np.array(st.spacing)) / st.spacing
            f = g - np.floor(g)
            if np.any(f < 1e-3) or np.any(f > 1 - 1e-3):
                continue                          # the kink at a cell face lies within h
            for d in range(3):
                def e_at(dx):
                    x2 = xyz.copy()
                    x2[a, d] += dx
                    one = [(mi[a:a + 1], q[a:a + 1], x2[a:a + 1])]
                    return _host_scores(st, one, maps)[0]
                fd = (e_at(h) - e_at(-h)) / (2 * h)
                assert abs(grad[o + a, d] - fd) <= 1e-3 * max(1.0, abs(fd)), (p, a, d, grad[o + a, d], fd)
        o += len(mi)


def test_adjoint_and_repeatability():
    print("[gpu] scorer adjoint", flush=True)
    st = _grid()
    rng = np.random.default_rng(3)
    maps = rng.normal(0, 2.0, size=(st.n_maps, st.points)).astype(np.float32)
    poses = _poses(st, n_poses=20, seed=9)
    gsc = rng.normal(0, 1.0, size=len(poses)).astype(np.float32)
    dmaps_n = st.n_maps * st.points
    dg, dd, dm, ds = Dev(4 * len(poses)), Dev(4 * dmaps_n), Dev(4 * dmaps_n), Dev(4 * len(poses))
    try:
        dg.put(gsc)
        dm.put(maps)
        dd.put(np.full(dmaps_n, 7.0, np.float32))    # overwritten everywhere
        with ag.Context(0) as ctx, ag.Scorer(ctx, st, poses) as sc:
            sc.adjoint(dg.p.value, dd.p.value)
            a1 = dd.get(dmaps_n)
            sc.adjoint(dg.p.value, dd.p.value)
            a2 = dd.get(dmaps_n)
            sc.score(dm.p.value, ds.p.value)
            s = ds.get(len(poses)).astype(np.float64)
    finally:
        dg.free(); dd.free(); dm.free(); ds.free()
    assert np.array_equal(a1, a2)
    ref = ps.cotangent(ps.splat(st, poses), dmaps_n, gsc.astype(np.float64))
    assert np.allclose(a1, ref, rtol=1e-5, atol=1e-6), np.max(np.abs(a1 - ref))
    lhs = float(a1.astype(np.float64) @ maps.reshape(-1).astype(np.float64))
    rhs = float(gsc.astype(np.float64) @ s)
    assert abs(lhs - rhs) <= 1e-5 * max(1.0, abs(rhs)), (lhs, rhs)


def test_outside_atoms():
    print("[gpu] scorer outside atoms", flush=True)
    st = _grid()
    maps = np.ones((st.n_maps, st.points), np.float32)
    hi = np.array([(st.n1[k] - 1 - st.ne[k]) * st.spacing for k in range(3)])
    lo = np.array([-st.ne[k] * st.spacing for k in range(3)])
    xyz = np.array([lo - 0.01, hi, hi - 0.01, lo])   # out, out (on the far face), in, in
    poses = [(np.array([0, 0, 0, 0]), np.zeros(4), xyz)]
    got, grad = _run(st, poses, maps)
    assert got[0] == pytest.approx(2 * 123.5 + 2 * 1.0, rel=1e-6)
    assert np.all(grad[:2] == 0.0)
