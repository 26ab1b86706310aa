"""The batched and differentiable workloads at their full shapes.

Batched ensemble: 64 receptor conformations (seed 1005) of 1,594-11,710
atoms with cubic boxes of 41-155 points, maps A C HD N NA OA + e + d.
  * the graph path (ag_batch_create, then ag_plan_execute, and
    ag_plan_execute_host into pinned memory) and the chunked path with
    pinned host outputs (ag_maps_batch) each give every
    conformation's maps bit for bit equal to ag_maps_plain's, at the plain
    path's 4 A cells (the same per-point summation order);
  * the autotuned batch is repeatable bit for bit and matches the FP64
    oracle, at every point of sampled z-planes of the smallest and the
    largest conformation, under the map tolerance.

Differentiable mode: the 5,000-atom full-workload receptor on a 65^3 box
(npts 64) with maps A C HD N OA + e + d, and the cotangent of 256 seeded
ligand poses of 20-60 atoms (types C OA N HD A, charges N(0, 0.15 e)),
L = sum of their trilinearly interpolated scores.
  * the analytic forward against the oracle's FP64 analytic model at 4,000
    sampled points, under the map tolerance;
  * the adjoint against FP64 central finite differences of the oracle's
    model (h = 1e-4 A for coordinates and radii, 1e-5 e for charges, 1e-5
    relative for eps and the weights) for coordinates of the atoms with the
    largest gradients, of carbons whose only role is to orient a bonded
    H-bonding atom's direction vectors (their gradient flows through the
    bond-vector adjoint), charges, Rij / eps and the four weights.  Each
    component must satisfy |g - fd| <= 1e-4 |fd| + 1e-4 max|fd| of its
    group.  A component that misses the bound at h is evaluated again at
    h/2; if the two differences disagree a kink of the model lies within h
    and the component is reported and skipped (a small minority at most).
"""
from __future__ import annotations

import ctypes as C
import os
import time

import numpy as np
import pytest

import autogrid_amd as ag
import oracle_lib as ol
import parity
from autogrid_amd import poses as ps
from autogrid_amd import synth

pytestmark = pytest.mark.gpu

PLAIN_TUNE = {"cell_size": 4.0, "cube": True}


def _hip():
    return C.CDLL("libamdhip64.so.7")


class Pinned:
    """A pinned host buffer of `nbytes` (hipHostMalloc)."""

    def __init__(self, nbytes: int):
        self.hip = _hip()
        self.p = C.c_void_p()
        assert self.hip.hipHostMalloc(C.byref(self.p), C.c_size_t(nbytes), 0) == 0
        self.raw = np.ctypeslib.as_array(C.cast(self.p, C.POINTER(C.c_uint8)), shape=(nbytes,))

    def free(self):
        self.hip.hipHostFree(self.p)


# ---------------------------------------------------------------- ensemble

@pytest.fixture(scope="module")
def ensemble64(tmp_path_factory):
    ens = synth.make_ensemble(str(tmp_path_factory.mktemp("ens64")), n_conf=64, seed=1005)
    states = [synth.conformation_state(ens, i) for i in range(len(ens.confs))]
    n_atoms = [st.n_atoms for st in states]
    npts = [st.n1[0] - 1 for st in states]
    assert len(states) == 64
    assert 1500 <= min(n_atoms) and max(n_atoms) <= 12000 and 40 <= min(npts) and max(npts) <= 160
    print(f"[gpu] ensemble: atoms {min(n_atoms)}-{max(n_atoms)}, npts {min(npts)}-{max(npts)}, "
          f"{sum(st.points for st in states)} points", flush=True)
    return states


@pytest.fixture(scope="module")
def plain64(ensemble64):
    """Every conformation's maps from 64 ag_maps_plain calls."""
    with ag.Context(0) as ctx:
        return [ctx.maps_plain(st) for st in ensemble64]


def test_ensemble64_graph_path_equals_plain(ensemble64, plain64):
    print("[gpu] 64-conformation batch, one graph", flush=True)
    hip = _hip()
    with ag.Context(0) as ctx, ctx.batch(ensemble64, tune=PLAIN_TUNE) as b:
        nbytes = 4 * b.map_floats
        d = C.c_void_p()
        assert hip.hipMalloc(C.byref(d), C.c_size_t(nbytes)) == 0
        pinned = Pinned(nbytes)
        try:
            b.execute(d.value)
            b.execute(d.value)             # a replay of the captured graph
            flat = np.empty(b.map_floats, np.float32)
            assert hip.hipMemcpy(flat.ctypes.data_as(C.c_void_p), d, C.c_size_t(nbytes), 2) == 0
            # the same graph with its host copies overlapping the far kernel
            host = pinned.raw.view(np.float32)
            host[:] = np.nan
            outs, o = [], 0
            for st, ref in zip(ensemble64, plain64):
                outs.append(host[o:o + ref.size].reshape(ref.shape))
                o += ref.size
            b.execute_host(d.value, outs)
            for i, (g, ref) in enumerate(zip(outs, plain64)):
                assert np.array_equal(g.view(np.uint32), ref.view(np.uint32)), f"host copies, conformation {i}"
        finally:
            hip.hipFree(d)
            pinned.free()
        offs = b.map_offsets
    for i, (st, ref) in enumerate(zip(ensemble64, plain64)):
        got = flat[offs[i]:offs[i] + ref.size]
        assert np.array_equal(got.view(np.uint32), ref.reshape(-1).view(np.uint32)), f"conformation {i}"


def test_ensemble64_pinned_chunks_equal_plain(ensemble64, plain64):
    print("[gpu] 64-conformation batch, pinned chunked outputs", flush=True)
    sizes = [ref.size for ref in plain64]
    buf = Pinned(4 * sum(sizes))
    try:
        flat = buf.raw.view(np.float32)
        flat[:] = np.nan
        outs, o = [], 0
        for st, n in zip(ensemble64, sizes):
            outs.append(flat[o:o + n].reshape(st.n_maps, st.n1[2], st.n1[1], st.n1[0]))
            o += n
        with ag.Context(0) as ctx:
            ctx.maps_batch(ensemble64, tune=PLAIN_TUNE, outs=outs)
        for i, (g, ref) in enumerate(zip(outs, plain64)):
            assert np.array_equal(g.view(np.uint32), ref.view(np.uint32)), f"conformation {i}"
    finally:
        buf.free()


def test_ensemble64_autotuned_against_oracle(ensemble64):
    print("[gpu] 64-conformation batch, autotuned, against the oracle", flush=True)
    with ag.Context(0) as ctx:
        with ctx.batch(ensemble64) as b:
            tune = b.tune()
        print("tune", tune, flush=True)
        got = ctx.maps_batch(ensemble64, tune=tune)
        again = ctx.maps_batch(ensemble64, tune=tune)
    for i, (g, h) in enumerate(zip(got, again)):
        assert np.array_equal(g.view(np.uint32), h.view(np.uint32)), f"repeat, conformation {i}"
    work = [st.points * st.n_atoms for st in ensemble64]
    small, large = int(np.argmin(work)), int(np.argmax(work))
    for i in (small, large):
        st = ensemble64[i]
        for z in ((0, st.n1[2] // 2, st.n1[2] - 1) if i == small else (st.n1[2] // 2,)):
            t0 = time.perf_counter()
            ref = ol.maps(st, z, z + 1)[:, z]
            print(f"oracle conformation {i} ({st.n_atoms} atoms, {st.n1[0]}^3) plane {z}: "
                  f"{time.perf_counter() - t0:.1f} s", flush=True)
            for k in range(st.n_maps):
                ok, mx, rms, bad = parity.within(got[i][k, z], ref[k])
                assert ok, f"conformation {i} map {k} plane {z}: max {mx}, {bad} points over tolerance"


# ------------------------------------------------------- analytic, 65^3

@pytest.fixture(scope="module")
def analytic65(tmp_path_factory):
    work = str(tmp_path_factory.mktemp("an65"))
    info = synth.make_workload("analytic", work)
    cwd = os.getcwd()
    os.chdir(work)
    try:
        st = ag.Job(info["gpf"]).prepare().state()
    finally:
        os.chdir(cwd)
    assert st.n_atoms == 5000 and st.n1 == (65, 65, 65)
    lig = {t: i for i, t in enumerate(synth.WORKLOADS["analytic"].ligand_types)}
    poses = ps.make_poses(st, lig, n_poses=256, seed=1005)
    assert len(poses) == 256 and all(20 <= len(p[0]) <= 60 for p in poses)
    G = ps.cotangent(ps.splat(st, poses), st.n_maps * st.points, np.ones(len(poses)))
    return st, G


class Dev:
    def __init__(self, nbytes):
        self.hip = _hip()
        self.p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.p), C.c_size_t(nbytes)) == 0

    def put(self, a):
        a = np.ascontiguousarray(a, np.float32)
        assert self.hip.hipMemcpy(self.p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0

    def get(self, n):
        out = np.empty(n, np.float32)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p, C.c_size_t(out.nbytes), 2) == 0
        return out

    def free(self):
        self.hip.hipFree(self.p)


def test_analytic65_forward_against_oracle(analytic65):
    print("[gpu] analytic forward, 5,000 atoms, 65^3", flush=True)
    st, _ = analytic65
    with ag.Context(0) as ctx, ag.Analytic(ctx, st) as an:
        d = Dev(4 * st.n_maps * st.points)
        try:
            an.forward(d.p.value)
            got = d.get(st.n_maps * st.points).reshape(st.n_maps, -1)
        finally:
            d.free()
    pts = np.sort(np.random.default_rng(11).choice(st.points, 4000, replace=False))
    ref = ol.analytic_maps(st, pts)
    for m in range(st.n_maps):
        ok, mx, rms, bad = parity.within(got[m, pts], ref[m])
        assert ok, f"map {m}: max {mx} rms {rms}, {bad} points over tolerance"


def _orienting_carbons(st, near_atoms, n=3):
    """Carbons bonded (<= 1.7 A) to an H-bonding atom among `near_atoms`:
    their coordinates enter that atom's direction vectors."""
    near = set(int(i) for i in near_atoms)
    hb = [int(i) for i in np.nonzero(np.isin(st.hbond, (4, 5)))[0] if int(i) in near]   # A1, A2 acceptors
    out = []
    for a in hb:
        lo, hi = max(0, a - 20), min(st.n_atoms, a + 21)
        for j in range(lo, hi):
            if j != a and st.hbond[j] == 0 and np.linalg.norm(st.coord[j] - st.coord[a]) <= 1.7 and j not in out:
                out.append(j)
                break
        if len(out) == n:
            break
    return out


def test_analytic65_adjoint_against_finite_differences(analytic65):
    print("[gpu] analytic adjoint vs FP64 FD, 5,000 atoms, 65^3, 256 poses", flush=True)
    st, G = analytic65
    support = np.nonzero(G)[0]
    pts = np.unique(support % st.points)
    Gs = G.reshape(st.n_maps, st.points)[:, pts]
    with ag.Context(0) as ctx, ag.Analytic(ctx, st) as an:
        dm, dg = Dev(4 * G.size), Dev(4 * G.size)
        try:
            an.forward(dm.p.value)
            dg.put(G)
            g = an.adjoint(dg.p.value)
        finally:
            dm.free()
            dg.free()
        rij0, eps0, w0 = an.rij0.copy(), an.eps0.copy(), an.weights0.copy()
    base = dict(coord=st.coord.copy(), charge=st.charge.copy(), rij=rij0, eps=eps0, weights=w0)

    def loss(**kw):
        p = {k: v.copy() for k, v in base.items()}
        p.update(kw)
        m = ol.analytic_maps(st, pts, rij=p["rij"], eps0=p["eps"], weights=p["weights"], coord=p["coord"],
                             charge=p["charge"])
        return float(np.sum(Gs * m))

    comps = []
    top = np.argsort(np.abs(g["coord"]).max(axis=1))[::-1]
    for i in top[:6]:
        comps += [("coord", (int(i), k), 1e-4) for k in range(3)]
    carbons = _orienting_carbons(st, top[:400])
    assert carbons, "no carbon orienting an acceptor near the poses"
    for i in carbons:
        comps += [("coord", (int(i), k), 1e-4) for k in range(3)]
    comps += [("charge", (int(i),), 1e-5) for i in np.argsort(np.abs(g["charge"]))[::-1][:4]]
    for name in ("rij", "eps"):
        flat = np.abs(g[name]).ravel()
        comps += [(name, tuple(int(x) for x in np.unravel_index(f, g[name].shape)), None)
                  for f in np.argsort(flat)[::-1][:3] if flat[f] > 0]
    comps += [("weights", (k,), None) for k in range(4)]

    t0 = time.perf_counter()
    rows = {}
    for name, idx, h in comps:
        v0 = base[name][idx]
        hh = h if h is not None else 1e-5 * max(abs(v0), 1e-3)

        def fd(step, name=name, idx=idx, v0=v0):
            a = base[name].copy()
            a[idx] = v0 + step
            lp = loss(**{name: a})
            a[idx] = v0 - step
            lm = loss(**{name: a})
            return (lp - lm) / (2 * step)
        rows.setdefault(name, []).append((idx, float(g[name][idx]), hh, fd(hh), fd))
    checked, skipped = 0, 0
    for name, rs in rows.items():
        scale = max(abs(r[3]) for r in rs)
        for idx, gv, hh, f1, fd in rs:
            if abs(gv - f1) <= 1e-4 * abs(f1) + 1e-4 * scale:
                checked += 1
                continue
            f2 = fd(hh / 2)
            if abs(f1 - f2) > 1e-5 * max(abs(f1), 1e-3 * scale):
                skipped += 1
                print("non-smooth FD", name, idx, gv, f1, f2, flush=True)
                continue
            checked += 1
            assert abs(gv - f2) <= 1e-4 * abs(f2) + 1e-4 * scale, (name, idx, gv, f1, f2, scale)
    print(f"FD: {checked} components checked, {skipped} skipped as non-smooth, "
          f"{time.perf_counter() - t0:.1f} s", flush=True)
    assert checked >= 35 and skipped <= checked // 5
