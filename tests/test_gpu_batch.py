"""Batched map sets (ag_batch_create / ag_maps_batch) on the GPU.

A batch runs many receptor conformations, with their own atoms and boxes,
through one replay of its captured graphs.  For a fixed near-kernel layout with the plain
path's 4 A cells every per-point sum runs in the same order as in a single
plan, so each conformation's maps must equal ag_maps_plain's bit for bit;
other cell sizes reorder the FP32 sums and are held to the oracle under the
parity tolerance.  Graph replays are checked against the first result.
"""
from __future__ import annotations

import dataclasses

import numpy as np
import pytest

import autogrid_amd as ag
import oracle_lib as ol
import parity
from autogrid_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ens(tmp_path_factory):
    d = tmp_path_factory.mktemp("ens")
    return synth.make_ensemble(str(d), n_conf=6, seed=3001, base_atoms=3000, n_range=(600, 3000),
                               npts_range=(16, 40))


@pytest.fixture(scope="module")
def ensemble(ens):
    return [synth.conformation_state(ens, i) for i in range(len(ens.confs))]


def _hip():
    import ctypes as C
    return C.CDLL("libamdhip64.so.7")


@pytest.mark.parametrize("cube", [False, True])
def test_batch_equals_plain_bitwise(ensemble, cube):
    print(f"[gpu] batch vs plain cube={cube}", flush=True)
    with ag.Context(0) as ctx:
        got = ctx.maps_batch(ensemble, tune={"cell_size": 4.0, "cube": cube})
        for st, g in zip(ensemble, got):
            ref = ctx.maps_plain(st)
            assert np.array_equal(g.view(np.uint32), ref.view(np.uint32))


def _wide_box(st, spacing=2.0, npts=80):
    """The conformation on a box of +-80 A: most of its pairs lie past
    64 A, and its corners reach past the far kernel's largest table."""
    return dataclasses.replace(st, spacing=spacing, n1=(npts + 1,) * 3, ne=(npts // 2,) * 3, _keep=[])


def test_batch_mixed_far_paths_equal_plain(ensemble):
    """A batch whose conformations take different far-kernel paths: small
    boxes go through the dielectric table atom by atom, wide boxes take the
    table-free path for atoms past bin 8192 of a wave's box and cap their
    table indices.  Each conformation's path depends on it alone, so every
    conformation still equals its own ag_maps_plain bit for bit; the wide
    boxes are also held to the oracle."""
    print("[gpu] batch with small and wide boxes vs plain", flush=True)
    sts = [ensemble[0], _wide_box(ensemble[1]), ensemble[2], _wide_box(ensemble[3])]
    wide = sts[1]
    # precondition: a large share of the wide box's pairs lies past 64 A,
    # and its corners pass bin 16380
    ax = (np.arange(8) + 0.5) / 8 * (wide.n1[0] - 1)
    pts = (np.stack(np.meshgrid(ax, ax, ax, indexing="ij"), -1).reshape(-1, 3) - wide.ne[0]) * wide.spacing
    r = np.linalg.norm(pts[:, None, :] - wide.coord[None, ::7, :], axis=-1)
    assert (r >= 64.0).mean() > 0.3
    corner = np.array([wide.ne[0] * wide.spacing] * 3)
    assert np.linalg.norm(np.abs(wide.coord) + corner, axis=1).max() * 128 > 16380
    with ag.Context(0) as ctx:
        got = ctx.maps_batch(sts, tune={"cell_size": 4.0, "cube": True})
        for i, (st, g) in enumerate(zip(sts, got)):
            ref = ctx.maps_plain(st)
            assert np.array_equal(g.view(np.uint32), ref.view(np.uint32)), f"conformation {i}"
    for st, g in ((sts[1], got[1]), (sts[3], got[3])):
        for z in (0, st.n1[2] // 2):
            ref = ol.maps(st, z, z + 1)[:, z]
            for k in range(st.n_maps):
                ok, mx, rms, bad = parity.within(g[k, z], ref[k])
                assert ok, f"map {k} plane {z}: max {mx}, {bad} points over tolerance"


def test_batch_other_cells_against_oracle(ensemble):
    print("[gpu] batch cell 3.0 vs oracle", flush=True)
    with ag.Context(0) as ctx:
        got = ctx.maps_batch(ensemble, tune={"cell_size": 3.0, "cube": True})
    for st, g in zip(ensemble, got):
        nz = st.n1[2]
        for z in (0, nz // 2, nz - 1):
            ref = ol.maps(st, z, z + 1)[:, z]
            for k in range(st.n_maps):
                ok, mx, rms, bad = parity.within(g[k, z], ref[k])
                assert ok, f"map {k} plane {z}: max {mx}, {bad} points over tolerance"


def test_batch_autotuned_and_graph_replays(ensemble):
    print("[gpu] batch autotune + replays", flush=True)
    import ctypes as C
    hip = _hip()
    with ag.Context(0) as ctx:
        with ctx.batch(ensemble) as b:
            tune = b.tune()
            print("tune", tune, flush=True)
            assert tune["cell_size"] in (3.0, 4.0, 5.0)
            nbytes = b.map_floats * 4
            d = C.c_void_p()
            assert hip.hipMalloc(C.byref(d), C.c_size_t(nbytes)) == 0
            try:
                b.execute(d.value)
                first = np.empty(b.map_floats, np.float32)
                assert hip.hipMemcpy(first.ctypes.data_as(C.c_void_p), d, C.c_size_t(nbytes), 2) == 0
                for rep in range(3):
                    # eager work between replays: a plain map set on the same context
                    ctx.maps_plain(ensemble[rep])
                    assert hip.hipMemset(d, 0, C.c_size_t(nbytes)) == 0
                    assert hip.hipDeviceSynchronize() == 0
                    b.execute(d.value)
                    again = np.empty_like(first)
                    assert hip.hipMemcpy(again.ctypes.data_as(C.c_void_p), d, C.c_size_t(nbytes), 2) == 0
                    assert np.array_equal(first.view(np.uint32), again.view(np.uint32)), f"replay {rep}"
            finally:
                hip.hipFree(d)
            offs = b.map_offsets
        # the graph's output equals the oracle within tolerance
        for i, st in enumerate(ensemble[:3]):
            g = first[offs[i]:offs[i] + st.n_maps * st.points].reshape(st.n_maps, st.n1[2], st.n1[1], st.n1[0])
            z = st.n1[2] // 3
            ref = ol.maps(st, z, z + 1)[:, z]
            for k in range(st.n_maps):
                assert parity.within(g[k, z], ref[k])[0]


def test_batch_distance_grids_and_vina(ens, ensemble):
    print("[gpu] batch fmap/cmap + vina", flush=True)
    sts = [dataclasses.replace(s, floating_grid=1, constriction_grid=1, constriction_cutoff=6.5, _keep=[])
           for s in ensemble[:3]]
    with ag.Context(0) as ctx:
        ctx.maps_batch(sts, tune={"cell_size": 4.0, "cube": False})
        fms, cms = ctx.last_fmaps, ctx.last_cmaps
        for st, fm, cm in zip(sts, fms, cms):
            ctx.maps_plain(st)
            assert np.array_equal(fm, ctx.last_fmap)
            assert np.array_equal(cm, ctx.last_cmap)
    # Vina tables of the base map set, shared by every conformation
    vt = ag.vina_tables(ensemble[0], ens.lig_types, ens.rec_types)
    vs = [dataclasses.replace(s, e_vdw_hb=vt, _keep=[]) for s in ensemble[:3]]
    with ag.Context(0) as ctx:
        got = ctx.maps_batch(vs, tune={"cell_size": 4.0, "cube": True}, potential=ag.AG_POTENTIAL_VINA)
        for st, g in zip(vs, got):
            assert np.array_equal(g.view(np.uint32), ctx.maps_vina(st).view(np.uint32))
    z = vs[0].n1[2] // 2
    ref = ol.maps_vina(vs[0], z, z + 1)[:, z]
    for k in range(vs[0].n_maps):
        assert parity.within(got[0][k, z], ref[k])[0]



@pytest.fixture(scope="module")
def ensemble12(tmp_path_factory):
    d = tmp_path_factory.mktemp("ens12")
    e = synth.make_ensemble(str(d), n_conf=12, seed=3002, base_atoms=2000, n_range=(400, 2000),
                            npts_range=(16, 48))
    return [synth.conformation_state(e, i) for i in range(len(e.confs))]


@pytest.mark.parametrize("which", ["ensemble", "ensemble12"])
def test_batch_pinned_outputs_pipeline(request, which):
    """With pinned host outputs ag_maps_batch runs its chunked pipeline
    (chunks alternating between two compute streams, each copied to the host
    while later chunks compute).  Twelve conformations give the full chunk
    pattern; six give fewer chunks than it asks for.  The maps must equal
    the pageable-output path's bit for bit, conformation by conformation."""
    import ctypes as C
    states = request.getfixturevalue(which)
    print(f"[gpu] pinned batch outputs, {len(states)} conformations", flush=True)
    hip = _hip()
    sizes = [st.n_maps * st.points for st in states]
    buf = C.c_void_p()
    assert hip.hipHostMalloc(C.byref(buf), C.c_size_t(4 * sum(sizes)), 0) == 0
    try:
        flat = np.ctypeslib.as_array(C.cast(buf, C.POINTER(C.c_float)), shape=(sum(sizes),))
        flat[:] = np.nan
        outs, o = [], 0
        for st, n in zip(states, sizes):
            outs.append(flat[o:o + n].reshape(st.n_maps, st.n1[2], st.n1[1], st.n1[0]))
            o += n
        tune = {"cell_size": 4.0, "cube": True}
        with ag.Context(0) as ctx:
            ref = ctx.maps_batch(states, tune=tune)
            for rep in range(2):
                ctx.maps_batch(states, tune=tune, outs=outs)
                for i, (g, r) in enumerate(zip(outs, ref)):
                    assert np.array_equal(g.view(np.uint32), r.view(np.uint32)), f"rep {rep} conformation {i}"
                flat[:] = np.nan
    finally:
        hip.hipHostFree(buf)


def test_batch_pinned_outputs_with_distance_grids(ensemble12):
    """The chunked pipeline with floating and constriction grids: every
    destination (maps, fmap, cmap) in pinned host memory, so each chunk's
    fmap / cmap copies run beside its maps.  All three must equal the
    pageable path's bit for bit."""
    import ctypes as C
    states = [dataclasses.replace(s, floating_grid=1, constriction_grid=1, constriction_cutoff=6.5, _keep=[])
              for s in ensemble12]
    print(f"[gpu] pinned batch outputs with fmap/cmap, {len(states)} conformations", flush=True)
    hip = _hip()
    sizes = [st.n_maps * st.points for st in states]
    pts = [st.points for st in states]
    total = 4 * (sum(sizes) + 2 * sum(pts))
    buf = C.c_void_p()
    assert hip.hipHostMalloc(C.byref(buf), C.c_size_t(total), 0) == 0
    try:
        raw = np.ctypeslib.as_array(C.cast(buf, C.POINTER(C.c_uint8)), shape=(total,))
        outs, fms, cms, o = [], [], [], 0
        for st, n in zip(states, sizes):
            outs.append(raw[o:o + 4 * n].view(np.float32).reshape(st.n_maps, st.n1[2], st.n1[1], st.n1[0]))
            o += 4 * n
        for st in states:
            fms.append(raw[o:o + 4 * st.points].view(np.float32).reshape(st.n1[2], st.n1[1], st.n1[0]))
            o += 4 * st.points
        for st in states:
            cms.append(raw[o:o + 4 * st.points].view(np.int32).reshape(st.n1[2], st.n1[1], st.n1[0]))
            o += 4 * st.points
        raw[:] = 0xFF
        tune = {"cell_size": 4.0, "cube": True}
        with ag.Context(0) as ctx:
            ref = ctx.maps_batch(states, tune=tune)
            rf, rc = ctx.last_fmaps, ctx.last_cmaps
            ctx.maps_batch(states, tune=tune, outs=outs, fmaps=fms, cmaps=cms)
            for i in range(len(states)):
                assert np.array_equal(outs[i].view(np.uint32), ref[i].view(np.uint32)), f"maps {i}"
                assert np.array_equal(fms[i].view(np.uint32), rf[i].view(np.uint32)), f"fmap {i}"
                assert np.array_equal(cms[i], rc[i]), f"cmap {i}"
    finally:
        hip.hipHostFree(buf)


@pytest.mark.parametrize("bhtree", [0, 1])
def test_plan_execute_host_equals_device_outputs(ensemble12, bhtree):
    """ag_plan_execute_host: the batch's graphs, with every map but e (and
    cmap, and fmap unless the tree build's collision rule rewrites it after
    the far kernel) copied to the host behind the event recorded after the
    near-field graph, the rest after the far-field graph.  Maps, fmap and cmap
    must equal ag_plan_execute's device outputs bit for bit: into pinned
    destinations on two launches (the second a replay) and into pageable
    ones.  A single-conformation plan must equal ag_maps_plain."""
    import ctypes as C
    states = [dataclasses.replace(s, floating_grid=1, constriction_grid=1, constriction_cutoff=6.5, bhtree=bhtree,
                                  _keep=[]) for s in ensemble12]
    print(f"[gpu] plan execute to host, bhtree={bhtree}, {len(states)} conformations", flush=True)
    hip = _hip()
    sizes = [st.n_maps * st.points for st in states]
    pts = [st.points for st in states]
    total = 4 * (sum(sizes) + 2 * sum(pts))
    buf = C.c_void_p()
    assert hip.hipHostMalloc(C.byref(buf), C.c_size_t(total), 0) == 0
    dev = [C.c_void_p() for _ in range(3)]
    for d, n in zip(dev, (sum(sizes), sum(pts), sum(pts))):
        assert hip.hipMalloc(C.byref(d), C.c_size_t(4 * n)) == 0
    try:
        raw = np.ctypeslib.as_array(C.cast(buf, C.POINTER(C.c_uint8)), shape=(total,))
        outs, fms, cms, o = [], [], [], 0
        for st, n in zip(states, sizes):
            outs.append(raw[o:o + 4 * n].view(np.float32).reshape(st.n_maps, st.n1[2], st.n1[1], st.n1[0]))
            o += 4 * n
        for st in states:
            fms.append(raw[o:o + 4 * st.points].view(np.float32).reshape(st.n1[2], st.n1[1], st.n1[0]))
            o += 4 * st.points
        for st in states:
            cms.append(raw[o:o + 4 * st.points].view(np.int32).reshape(st.n1[2], st.n1[1], st.n1[0]))
            o += 4 * st.points
        with ag.Context(0) as ctx:
            with ctx.batch(states, tune={"cell_size": 4.0, "cube": True}) as b:
                b.execute(dev[0].value, dev[1].value, dev[2].value)
                ref = np.empty(sum(sizes), np.float32)
                rf = np.empty(sum(pts), np.float32)
                rc = np.empty(sum(pts), np.int32)
                for a, d in zip((ref, rf, rc), dev):
                    assert hip.hipMemcpy(a.ctypes.data_as(C.c_void_p), d, C.c_size_t(a.nbytes), 2) == 0
                mo, po = b.map_offsets, b.point_offsets

                def check(got_maps, got_f, got_c, what):
                    for i, st in enumerate(states):
                        r = ref[mo[i]:mo[i] + sizes[i]]
                        assert np.array_equal(got_maps[i].reshape(-1).view(np.uint32), r.view(np.uint32)), \
                            f"{what}: maps {i}"
                        assert np.array_equal(got_f[i].reshape(-1).view(np.uint32),
                                              rf[po[i]:po[i] + pts[i]].view(np.uint32)), f"{what}: fmap {i}"
                        assert np.array_equal(got_c[i].reshape(-1), rc[po[i]:po[i] + pts[i]]), f"{what}: cmap {i}"
                for rep in range(2):
                    raw[:] = 0xFF
                    b.execute_host(dev[0].value, outs, fms, cms, dev[1].value, dev[2].value)
                    check(outs, fms, cms, f"pinned, launch {rep}")
                got = b.execute_host(dev[0].value, fmap_dev_ptr=dev[1].value, cmap_dev_ptr=dev[2].value)
                check(got, b.last_fmaps, b.last_cmaps, "pageable")
            st = states[-1]
            ref1 = ctx.maps_plain(st)
            f1, c1 = ctx.last_fmap, ctx.last_cmap
            with ctx.plan(st) as p:
                got = p.execute_host(dev[0].value, fmap_dev_ptr=dev[1].value, cmap_dev_ptr=dev[2].value)
                assert np.array_equal(got[0].view(np.uint32), ref1.view(np.uint32))
                assert np.array_equal(p.last_fmaps[0].view(np.uint32), f1.view(np.uint32))
                assert np.array_equal(p.last_cmaps[0], c1)
    finally:
        for d in dev:
            hip.hipFree(d)
        hip.hipHostFree(buf)
