"""Benchmark of the grid-map path on one MI355X.

Workload (default "full"): a seeded 5,000-atom receptor in a 45 A sphere,
npts 126 (127^3 = 2,048,383 points), maps A C N NA OA S SA HD F Cl Br I P
plus e and d (15 maps), AD4.1_bound parameters, distance-dependent
dielectric.  One step computes the whole map set on the device from
device-resident inputs: cell list, near-field brick kernel, far-field
electrostatics kernel.  `value` is that map-set time in seconds.

Also reported on the same JSON line, each with an `*_is` label:
  roofline      the far-field electrostatics kernel (the dominant one):
                algorithmic FLOPs 16 per (point, atom) pair, achieved
                TFLOP/s from its HIP-event time on the engine's stream,
                against the MI355X FP32 vector peak.
  wall_excl_ascii_s / wall_with_ascii_output_s
                the map set from host inputs (plan build, H2D, kernels,
                D2H into pinned memory), and the autogrid4 drop-in end to
                end with its ASCII output.
  accuracy      the GPU maps against the reference autogrid4's .map text
                (every value) and the FP64 oracle (3 z-planes); W_A; the
                Vina-potential maps against the oracle (3 z-planes).
  map_set_roofline
                the whole map set's FP32 work, 16 W_E + F_A W_A FLOPs
                (W_A from the CPU leg), over its kernels' time.
  cpu_baseline  the reference autogrid4 (built from its sources with
                OpenMP) on the whole workload on this host's cores, and a
                labelled z-slab of the large workload scaled by point count.
  batched       64 receptor conformations (1,545-11,710 atoms, cubic boxes
                of 41-155 points, maps A C HD N NA OA + e + d) against 64
                sequential single-receptor runs (tables, bond vectors,
                ag_maps_plain with its uploads and copies), both without
                ASCII output: the captured graphs with their host copies
                (one-shot and resident replays), and the chunked
                pinned-output pipeline.
  large         the 20,000-atom, 257^3 map set (seconds per map set,
                kernel times, far-field pairs/s).
  protein       the reference's x1hpv receptor at protein density on a
                121^3 box with its CASE0 map set.
  vina          the full receptor and box with the Vina potential
                (calc_vina_potential's tables): map set and kernel times.
  analytic      the differentiable mode on the full receptor with a 65^3
                box: forward and forward+adjoint times (cotangent from 256
                ligand poses) and a 200-step Adam fit of the AD4 weights.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_PEAK_TFLOPS = 157.3          # MI355X FP32 vector, packed
FP32_PEAK_NONPACKED_TFLOPS = 78.6
FLOPS_PER_PAIR_E = 16


def hip_runtime():
    return C.CDLL("libamdhip64.so.7")


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or "unknown"


def cpu_threads() -> int:
    """OpenMP threads for the reference: every core this process may run
    on, up to the reference's own cap of 32 (main.cpp:207), and within
    OMP_NUM_THREADS where the host sets it (its share of a shared machine)."""
    n = min(len(os.sched_getaffinity(0)), 32)
    env = os.environ.get("OMP_NUM_THREADS")
    return max(1, min(n, int(env))) if env and env.isdigit() else n


def run_reference(workload: str, threads: int, npts=None):
    """The reference autogrid4 (built from its sources by oracle/Makefile)
    on a workload in a scratch directory: (seconds incl. ASCII output,
    directory holding its maps, the workload info).  None if not built."""
    from autogrid_amd import synth
    ref = os.path.join(ROOT, "oracle", "_ref", "autogrid4")
    if not os.path.exists(ref):
        return None
    d = tempfile.mkdtemp(prefix="agref_")
    info = synth.make_workload(workload, d, npts=npts)
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    t0 = time.perf_counter()
    p = subprocess.run([ref, "-p", info["gpf"], "-l", "ref.glg"], cwd=d, env=env, capture_output=True)
    t = time.perf_counter() - t0
    if p.returncode != 0:
        shutil.rmtree(d, ignore_errors=True)
        return None
    return t, d, info


def read_map_values(path: str):
    import numpy as np
    with open(path, "rb") as f:
        body = f.read().split(b"\n", 6)[6]
    return np.array(body.split(), dtype=np.float64)


def cpu_leg(st, gpu_maps, map_files, large: bool = True, vina=None) -> dict:
    """The reference on the host's cores (the CPU baseline) and the checks
    of the GPU map set against it and the FP64 oracle.

    cpu_baseline  the full workload run in full by the reference autogrid4
                  with OpenMP (wall incl. its ASCII output), and a labelled
                  z-slab of the large workload scaled by point count.
    accuracy      the GPU map set (device-resident run) against the
                  reference's .map text at every point of every map (the
                  text carries 3 decimals, so the bound adds 5e-4), and
                  against the FP64 oracle at every point of 3 z-planes;
                  W_A, the exact count of (point, atom) pairs within 8 A,
                  from the oracle (main.cpp:3036-3049)."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    threads = cpu_threads()
    out = {"cpu_baseline": None, "accuracy": None}
    r = run_reference("full", threads)
    if r is None:
        return out
    t, d, info = r
    P, N = st.points, st.n_atoms
    base = {"value": t, "unit": "s", "cores": threads, "kind": "reference", "cpu_model": cpu_model(),
            "host_cores_visible": len(os.sched_getaffinity(0)),
            "sample": f"the whole full workload ({N} atoms, {P} points, {st.n_maps} maps), "
                      f"OpenMP {threads} threads, wall incl. ASCII map output",
            "pair_evals_per_s": P * N / t}
    acc = {}
    try:
        worst, sq, cnt, over = 0.0, 0.0, 0, 0
        for k, fn in enumerate(map_files):
            ref = read_map_values(os.path.join(d, fn))
            g = gpu_maps[k].reshape(-1).astype(np.float64)
            assert ref.size == g.size, fn
            err = np.abs(g - ref)
            lim = np.maximum(1e-3, 1e-3 * np.abs(ref)) + 5e-4
            worst = max(worst, float(err.max()))
            sq += float(np.sum(err ** 2))
            cnt += err.size
            over += int(np.count_nonzero(err > lim))
        acc["vs_reference_maps"] = {"max_abs": worst, "rms": (sq / cnt) ** 0.5, "values": cnt,
                                    "over_tolerance": over,
                                    "tolerance": "max(1e-3, 1e-3 |ref|) + 5e-4 (3-decimal text)"}
    finally:
        shutil.rmtree(d, ignore_errors=True)
    nz = st.n1[2]
    worst, sq, cnt, over = 0.0, 0.0, 0, 0
    for z in (0, nz // 2, nz - 1):
        ref = ol.maps(st, z, z + 1)[:, z].reshape(st.n_maps, -1)
        g = gpu_maps[:, z].reshape(st.n_maps, -1).astype(np.float64)
        err = np.abs(g - ref)
        worst = max(worst, float(err.max()))
        sq += float(np.sum(err ** 2))
        cnt += err.size
        over += int(np.count_nonzero(err > np.maximum(1e-3, 1e-3 * np.abs(ref))))
    acc["vs_fp64_oracle"] = {"max_abs": worst, "rms": (sq / cnt) ** 0.5, "values": cnt, "over_tolerance": over,
                             "planes": [0, nz // 2, nz - 1], "tolerance": "max(1e-3, 1e-3 |ref|)"}
    acc["W_A"] = ol.count_pairs(st)
    acc["W_E"] = P * N
    if vina is not None:
        # the Vina-potential map set of the same box against the oracle's
        # Vina tables and sums, at every point of 3 z-planes
        vst, vmaps = vina
        worst, sq, cnt, over = 0.0, 0.0, 0, 0
        for z in (0, nz // 2, nz - 1):
            ref = ol.maps_vina(vst, z, z + 1)[:, z].reshape(vst.n_maps, -1)
            g = vmaps[:, z].reshape(vst.n_maps, -1).astype(np.float64)
            err = np.abs(g - ref)
            worst = max(worst, float(err.max()))
            sq += float(np.sum(err ** 2))
            cnt += err.size
            over += int(np.count_nonzero(err > np.maximum(1e-3, 1e-3 * np.abs(ref))))
        acc["vina_vs_fp64_oracle"] = {"max_abs": worst, "rms": (sq / cnt) ** 0.5, "values": cnt,
                                      "over_tolerance": over, "planes": [0, nz // 2, nz - 1],
                                      "tolerance": "max(1e-3, 1e-3 |ref|)"}
    out["accuracy"] = acc
    if large:
        from autogrid_amd import synth
        w = synth.WORKLOADS["large"]
        # the reference's per-point work is the same at every point (both
        # loops run over every atom), so a centred sub-box of 3 planes
        # (npts 256 x 96 x 2, about 20 s on 16 cores) scales by point count
        box = (w.npts, 96, 2)
        r = run_reference("large", threads, npts=box)
        if r is not None:
            t3, d3, _ = r
            shutil.rmtree(d3, ignore_errors=True)
            n1 = w.npts + 1
            sample, full = (box[0] + 1) * (box[1] + 1) * (box[2] + 1), n1 ** 3
            base["large_extrapolated"] = {
                "value": t3 * full / sample, "unit": "s", "extrapolated": True,
                "sample": f"npts {box[0]} {box[1]} {box[2]} centred ({sample} points, {t3:.2f} s incl. ASCII output), "
                          f"scaled x{full / sample:.1f} by point count"}
    out["cpu_baseline"] = base
    return out


TRAFFIC_PROFILE = "profiles/full_mapset_5000atoms_127cubed.md"


def profiled_traffic():
    """HBM bytes per k_far launch on the full workload, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE runs (calibrated by a device copy);
    None when the summary is missing."""
    import re
    try:
        with open(os.path.join(ROOT, TRAFFIC_PROFILE)) as f:
            m = re.search(r"traffic_bytes_per_launch k_far: (\d+)", f.read())
    except OSError:
        return None
    return int(m.group(1)) if m else None


def batched_leg(reps: int = 2) -> dict:
    """The 64-conformation ensemble: batched against sequential single runs."""
    import numpy as np
    import autogrid_amd as ag
    from autogrid_amd import synth

    work = tempfile.mkdtemp(prefix="agens_")
    try:
        ens = synth.make_ensemble(work, n_conf=64, seed=1005)
    finally:
        shutil.rmtree(work, ignore_errors=True)
    inputs = [synth.conformation_state(ens, i, bonds=False) for i in range(len(ens.confs))]
    n_maps = inputs[0].n_maps
    sizes = [n_maps * st.points for st in inputs]
    total = sum(sizes)
    hip = hip_runtime()
    hbuf = C.c_void_p()
    if hip.hipHostMalloc(C.byref(hbuf), C.c_size_t(4 * total), 0) != 0:
        raise RuntimeError("hipHostMalloc failed")
    try:
        flat = np.ctypeslib.as_array(C.cast(hbuf, C.POINTER(C.c_float)), shape=(total,))
        outs, o = [], 0
        for st, n in zip(inputs, sizes):
            outs.append(flat[o:o + n].reshape(n_maps, st.n1[2], st.n1[1], st.n1[0]))
            o += n
        import dataclasses

        def with_bonds(st, tables):
            rexp, rv, rv2, dis = ag.bond_vectors(st, st.coord, st.type, st.hbond)
            ev, sol, reps_ = tables
            return dataclasses.replace(st, rexp=rexp, rvector=rv, rvector2=rv2, disorder=dis, e_vdw_hb=ev,
                                       sol_fn=sol, r_eps=reps_, _keep=[])

        with ag.Context(0) as ctx:
            def sequential():
                for st, out in zip(inputs, outs):
                    tables = ag.build_tables(st)
                    ctx.maps_plain(with_bonds(st, tables), out=out)

            def batch_states():
                ev, sol, reps_ = ag.build_tables(inputs[0])
                states = ag.bond_vectors_batch(inputs)
                return [dataclasses.replace(st, e_vdw_hb=ev, sol_fn=sol, r_eps=reps_, _keep=[]) for st in states]

            def batched(tune):
                ctx.maps_batch(batch_states(), tune=tune, outs=outs)

            def graph_once(tune):
                # host precompute, ag_batch_create, one launch of the captured
                # graphs (cell list and near pass, then the persistent far
                # kernel) with the maps other than e copied to pinned memory
                # during the far kernel and the e maps after it
                with ctx.batch(batch_states(), tune=tune) as b:
                    assert b.map_floats == total
                    b.execute_host(dmaps.value, outs)

            # autotune once (the choice is what a tuning cache would hold)
            t0 = time.perf_counter()
            tables0 = ag.build_tables(inputs[0])
            with ctx.batch([with_bonds(st, tables0) for st in inputs]) as b:
                tune = b.tune()
            tune_wall = time.perf_counter() - t0
            dmaps = C.c_void_p()
            if hip.hipMalloc(C.byref(dmaps), C.c_size_t(4 * total)) != 0:
                raise RuntimeError("hipMalloc failed")
            try:
                sequential()
                seq_copy = flat.copy()

                def diffs():
                    d = np.abs(flat - seq_copy)
                    return float(np.max(d)), float(np.max(d / np.maximum(1.0, np.abs(seq_copy))))
                batched(tune)
                diff, rel = diffs()
                graph_once(tune)
                gdiff, grel = diffs()
                t_seq, t_bat, t_graph, t_replay = [], [], [], []
                for _ in range(reps):
                    t1 = time.perf_counter()
                    sequential()
                    t_seq.append(time.perf_counter() - t1)
                    t1 = time.perf_counter()
                    batched(tune)
                    t_bat.append(time.perf_counter() - t1)
                    t1 = time.perf_counter()
                    graph_once(tune)
                    t_graph.append(time.perf_counter() - t1)
                # replays of one resident plan: graph launch with its copies
                with ctx.batch(batch_states(), tune=tune) as b:
                    b.execute(dmaps.value)
                    for _ in range(reps + 1):
                        t1 = time.perf_counter()
                        b.execute_host(dmaps.value, outs)
                        t_replay.append(time.perf_counter() - t1)
                    replay_gpu_ms = ctx.timing()["total_ms"]
            finally:
                hip.hipFree(dmaps)
    finally:
        hip.hipHostFree(hbuf)
    pairs = sum(st.points * st.n_atoms for st in inputs)
    tb, ts, tg, tr = min(t_bat), min(t_seq), min(t_graph), min(t_replay)
    return {"conformations": len(inputs), "points": sum(st.points for st in inputs), "pair_evals": pairs,
            "atoms_range": [min(st.n_atoms for st in inputs), max(st.n_atoms for st in inputs)],
            "npts_range": [min(st.n1[0] - 1 for st in inputs), max(st.n1[0] - 1 for st in inputs)],
            "sequential_s": ts,
            "sequential_is": "64 x (host tables + bond vectors + ag_maps_plain: plan build, H2D, kernels, "
                             "D2H into pinned memory), no ASCII",
            "graph_s": tg, "graph_speedup": ts / tg,
            "graph_is": "host tables + bond vectors for all conformations, ag_batch_create, ag_plan_execute_host: "
                        "the captured graphs (cell list and near pass, then the persistent far kernel), the maps other "
                        "than e copied into pinned memory during the far kernel, the e maps after it",
            "graph_replay_s": tr, "graph_replay_speedup": ts / tr, "graph_replay_gpu_ms": replay_gpu_ms,
            "graph_replay_is": "ag_plan_execute_host replay of the resident batch plan's graphs with their copies "
                               "into pinned memory",
            "chunked_s": tb, "chunked_speedup": ts / tb,
            "chunked_is": "ag_maps_batch with pinned outputs: host precompute, 10 chunks on two compute streams, "
                          "each chunk's D2H overlapping later chunks",
            "speedup": ts / tb, "batched_s": tb,
            "batched_pair_evals_per_s": pairs / tb,
            "tune": {"cell_size": tune["cell_size"], "cube": tune["cube"]}, "tune_ms": tune["tune_ms"],
            "tune_wall_s": tune_wall, "max_abs_diff_vs_sequential": diff, "max_rel_diff_vs_sequential": rel,
            "graph_max_rel_diff_vs_sequential": grel, "graph_max_abs_diff_vs_sequential": gdiff,
            "reps": reps}


def large_leg(reps: int = 3) -> dict:
    """The large blind-docking box: 20,000 atoms in an 80 A sphere, 257^3
    points, 15 maps, from device-resident inputs (as the headline line)."""
    import autogrid_amd as ag
    from autogrid_amd import synth

    work = tempfile.mkdtemp(prefix="aglarge_")
    try:
        info = synth.make_workload("large", work)
        cwd = os.getcwd()
        os.chdir(work)
        try:
            st = ag.Job(info["gpf"]).prepare().state()
        finally:
            os.chdir(cwd)
    finally:
        shutil.rmtree(work, ignore_errors=True)
    hip = hip_runtime()
    dptr = C.c_void_p()
    if hip.hipMalloc(C.byref(dptr), C.c_size_t(4 * st.n_maps * st.points)) != 0:
        raise RuntimeError("hipMalloc failed")
    try:
        with ag.Context(0) as ctx, ctx.plan(st) as plan:
            plan.execute(dptr.value)
            wall, tms = [], []
            for _ in range(reps):
                t0 = time.perf_counter()
                plan.execute(dptr.value)
                wall.append(time.perf_counter() - t0)
                tms.append(ctx.timing())
            far_pairs = plan.audit()["far_pairs"]    # outside the timed region
    finally:
        hip.hipFree(dptr)
    k = min(range(reps), key=lambda i: wall[i])
    far = tms[k]["far_ms"]
    return {"workload": f"large: {st.n_atoms} atoms, {st.n1[0]}^3 points, {st.n_maps} maps",
            "map_set_s": wall[k], "kernels_ms": {"cell_list": tms[k]["cell_ms"], "near": tms[k]["near_ms"],
                                               "far": far},
            "far_pairs": far_pairs,
            "pair_evals_per_s": far_pairs / (far * 1e-3),
            "far_tflops": FLOPS_PER_PAIR_E * far_pairs / (far * 1e-3) / 1e12}


def protein_leg(reps: int = 5) -> dict:
    """The reference's own receptor at protein density: x1hpv (HIV-1
    protease, 1,849 atoms, tests/golden/reftests) with its CASE0 map set
    (11 affinity maps + e + d) on a 121^3 box at 0.375 A around its grid
    centre, from device-resident inputs."""
    import autogrid_amd as ag

    src = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden", "reftests")
    work = tempfile.mkdtemp(prefix="agprot_")
    try:
        shutil.copy(os.path.join(src, "x1hpv.pdbqt"), work)
        gpf = open(os.path.join(src, "x1hpv_CASE0.gpf")).read().replace("npts 6 6 6", "npts 120 120 120")
        open(os.path.join(work, "x1hpv.gpf"), "w").write(gpf)
        cwd = os.getcwd()
        os.chdir(work)
        try:
            st = ag.Job("x1hpv.gpf").prepare().state()
        finally:
            os.chdir(cwd)
    finally:
        shutil.rmtree(work, ignore_errors=True)
    hip = hip_runtime()
    dptr = C.c_void_p()
    if hip.hipMalloc(C.byref(dptr), C.c_size_t(4 * st.n_maps * st.points)) != 0:
        raise RuntimeError("hipMalloc failed")
    try:
        with ag.Context(0) as ctx, ctx.plan(st) as plan:
            plan.execute(dptr.value)
            wall, tms = [], []
            for _ in range(reps):
                t0 = time.perf_counter()
                plan.execute(dptr.value)
                wall.append(time.perf_counter() - t0)
                tms.append(ctx.timing())
    finally:
        hip.hipFree(dptr)
    k = min(range(reps), key=lambda i: wall[i])
    return {"workload": f"x1hpv: {st.n_atoms} atoms, {st.n1[0]}^3 points, {st.n_maps} maps (CASE0 map set)",
            "map_set_s": wall[k], "kernels_ms": {"cell_list": tms[k]["cell_ms"], "near": tms[k]["near_ms"],
                                               "far": tms[k]["far_ms"]}}


def vina_leg(reps: int = 10):
    """The full workload's receptor and 127^3 box with the Vina potential
    (calc_vina_potential, main.cpp:1975-2127): Vina tables from the front
    end (ag_job_prepare_vina), the same 13 affinity maps, from
    device-resident inputs.  The e and d maps are zero under this potential,
    so the map set is the cell list and the near kernel.  Returns the leg's
    figures, the state and the maps (for the accuracy check of the CPU leg)."""
    import numpy as np

    import autogrid_amd as ag
    from autogrid_amd import synth

    work = tempfile.mkdtemp(prefix="agvina_")
    try:
        info = synth.make_workload("full", work)
        cwd = os.getcwd()
        os.chdir(work)
        try:
            st = ag.Job(info["gpf"]).prepare_vina().state()
        finally:
            os.chdir(cwd)
    finally:
        shutil.rmtree(work, ignore_errors=True)
    hip = hip_runtime()
    nbytes = 4 * st.n_maps * st.points
    dptr = C.c_void_p()
    if hip.hipMalloc(C.byref(dptr), C.c_size_t(nbytes)) != 0:
        raise RuntimeError("hipMalloc failed")
    try:
        with ag.Context(0) as ctx, ctx.plan(st, potential=ag.AG_POTENTIAL_VINA) as plan:
            plan.execute(dptr.value)
            wall, tms = [], []
            for _ in range(reps):
                t0 = time.perf_counter()
                plan.execute(dptr.value)
                wall.append(time.perf_counter() - t0)
                tms.append(ctx.timing())
        maps = np.empty((st.n_maps, st.n1[2], st.n1[1], st.n1[0]), np.float32)
        if hip.hipMemcpy(maps.ctypes.data_as(C.c_void_p), dptr, C.c_size_t(nbytes), 2) != 0:
            raise RuntimeError("hipMemcpy failed")
    finally:
        hip.hipFree(dptr)
    k = min(range(reps), key=lambda i: wall[i])
    near = sum(t["near_ms"] for t in tms) / reps
    leg = {"workload": f"full with Vina tables: {st.n_atoms} atoms, {st.n1[0]}^3 points, {st.n_maps} maps",
           "map_set_s": wall[k], "kernels_ms": {"cell_list": tms[k]["cell_ms"], "near": near},
           "pair_evals_per_s": st.points * st.n_atoms / (wall[k]),
           "pair_evals_per_s_is": "all P*N (point, atom) pairs of the map set over its wall time"}
    return leg, st, maps


def analytic_leg(reps: int = 10, fit_steps: int = 200) -> dict:
    """Analytic mode on the full workload's receptor with a 65^3 box (maps
    A C HD N OA + e + d): forward and forward+adjoint times with the
    cotangent of 256 seeded ligand poses' summed trilinear scores, and the
    200-step Adam fit of the four weights (torch, in a child process)."""
    import numpy as np

    import autogrid_amd as ag
    from autogrid_amd import poses as ps
    from autogrid_amd import synth

    work = tempfile.mkdtemp(prefix="agan_")
    try:
        info = synth.make_workload("analytic", work)
        cwd = os.getcwd()
        os.chdir(work)
        try:
            st = ag.Job(info["gpf"]).prepare().state()
        finally:
            os.chdir(cwd)
        lig = {t: i for i, t in enumerate(synth.WORKLOADS["analytic"].ligand_types)}
        poses = ps.make_poses(st, lig, n_poses=256, seed=1005)
        S = ps.splat(st, poses)
        cot = ps.cotangent(S, st.n_maps * st.points, np.ones(len(poses))).astype(np.float32)
        hip = hip_runtime()
        nbytes = 4 * st.n_maps * st.points
        dm, dg = C.c_void_p(), C.c_void_p()
        assert hip.hipMalloc(C.byref(dm), C.c_size_t(nbytes)) == 0
        assert hip.hipMalloc(C.byref(dg), C.c_size_t(nbytes)) == 0
        assert hip.hipMemcpy(dg, cot.ctypes.data_as(C.c_void_p), C.c_size_t(nbytes), 1) == 0
        ds = C.c_void_p()
        assert hip.hipMalloc(C.byref(ds), C.c_size_t(4 * len(poses))) == 0
        ones = np.ones(len(poses), np.float32)
        assert hip.hipMemcpy(ds, ones.ctypes.data_as(C.c_void_p), C.c_size_t(ones.nbytes), 1) == 0
        with ag.Context(0) as ctx, ag.Analytic(ctx, st) as an, ag.Scorer(ctx, st, poses) as sc:
            an.forward(dm.value)
            an.adjoint(dg.value)
            # the trilinear scoring kernel and its adjoint on the same maps
            ts, ta = [], []
            for _ in range(reps):
                t1 = time.perf_counter()
                sc.score(dm.value, ds.value)
                t2 = time.perf_counter()
                sc.adjoint(ds.value, dg.value)
                ts.append(t2 - t1)
                ta.append(time.perf_counter() - t2)
            assert hip.hipMemcpy(dg, cot.ctypes.data_as(C.c_void_p), C.c_size_t(nbytes), 1) == 0
            fw, fa, fk, ak = [], [], [], []
            for _ in range(reps):
                hip.hipDeviceSynchronize()
                t0 = time.perf_counter()
                an.forward(dm.value)
                t1 = time.perf_counter()
                fk.append(ctx.timing()["total_ms"])
                an.adjoint(dg.value)
                t2 = time.perf_counter()
                ak.append(ctx.timing()["total_ms"])
                fw.append(t1 - t0)
                fa.append(t2 - t0)
        hip.hipFree(dm)
        hip.hipFree(dg)
        hip.hipFree(ds)
        out = {"workload": f"analytic: {st.n_atoms} atoms, {st.n1[0]}^3 points, {st.n_maps} maps, 256 poses",
               "forward_s": min(fw), "forward_adjoint_s": min(fa), "ratio": min(fa) / min(fw),
               "forward_kernels_ms": min(fk), "adjoint_kernels_ms": min(ak),
               "kernel_ratio": (min(fk) + min(ak)) / min(fk),
               "score_poses_ms": 1e3 * min(ts), "score_adjoint_ms": 1e3 * min(ta)}
        if fit_steps:
            p = subprocess.run([sys.executable, "-m", "autogrid_amd.fit", "--workload", "analytic", "--steps",
                                str(fit_steps), "--poses", "256", "--workdir", work], cwd=ROOT,
                               capture_output=True, text=True, timeout=900)
            if p.returncode == 0:
                fit = json.loads(p.stdout.strip().splitlines()[-1])
                out["fit"] = {k: fit[k] for k in ("steps", "loss_initial", "loss_final", "w_rel_err_max",
                                                   "s_per_step", "scale_true")}
            else:
                out["fit"] = {"error": p.stderr[-400:]}
        return out
    finally:
        shutil.rmtree(work, ignore_errors=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="full", choices=["small", "full", "large"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-cli", action="store_true", help="skip the drop-in run with ASCII output")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-conformation batched comparison")
    ap.add_argument("--no-analytic", action="store_true", help="skip the analytic forward/adjoint leg")
    ap.add_argument("--no-large", action="store_true", help="skip the 20,000-atom, 257^3 map set")
    ap.add_argument("--no-protein", action="store_true", help="skip the x1hpv receptor at protein density")
    ap.add_argument("--no-vina", action="store_true", help="skip the Vina-potential map set of the full box")
    ap.add_argument("--fit-steps", type=int, default=200, help="Adam steps of the analytic leg's fit (0: none)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run one device-to-device copy of BYTES (counter calibration) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        hip = hip_runtime()
        a, b = C.c_void_p(), C.c_void_p()
        n = C.c_size_t(args.copy_calibration)
        assert hip.hipMalloc(C.byref(a), n) == 0 and hip.hipMalloc(C.byref(b), n) == 0
        assert hip.hipMemset(a, 1, n) == 0
        hip.hipDeviceSynchronize()
        assert hip.hipMemcpy(b, a, n, 3) == 0   # device to device
        hip.hipDeviceSynchronize()
        hip.hipFree(a)
        hip.hipFree(b)
        print(json.dumps({"copy_bytes": args.copy_calibration}))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (a map set would shard by "
              "z-slab across GPUs without communication).", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import autogrid_amd as ag
    from autogrid_amd import synth

    work = tempfile.mkdtemp(prefix="agbench_")
    try:
        info = synth.make_workload(args.workload, work)
        cwd = os.getcwd()
        os.chdir(work)
        try:
            t0 = time.perf_counter()
            st = ag.Job(info["gpf"]).prepare().state()
            frontend_s = time.perf_counter() - t0     # GPF + PDBQT read, tables, bond vectors
        finally:
            os.chdir(cwd)
        hip = hip_runtime()
        ctx = ag.Context(0)
        plan = ctx.plan(st)
        nbytes = st.n_maps * st.points * 4
        dptr = C.c_void_p()
        if hip.hipMalloc(C.byref(dptr), C.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
        for _ in range(args.warmup):
            plan.execute(dptr.value)
        hip.hipDeviceSynchronize()
        far_ms, near_ms, cell_ms = [], [], []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            plan.execute(dptr.value)          # synchronous: ends with a stream sync
            tm = ctx.timing()
            far_ms.append(tm["far_ms"])
            near_ms.append(tm["near_ms"])
            cell_ms.append(tm["cell_ms"])
        hip.hipDeviceSynchronize()
        t = time.perf_counter() - t0
        ms_per_step = 1000.0 * t / args.steps
        far_avg = sum(far_ms) / len(far_ms)
        P, N = st.points, st.n_atoms
        # outside the timed region: the maps for the accuracy checks, and the
        # table-bin audit against FP64 distances
        gpu_maps = np.empty((st.n_maps, st.n1[2], st.n1[1], st.n1[0]), np.float32)
        if hip.hipMemcpy(gpu_maps.ctypes.data_as(C.c_void_p), dptr, C.c_size_t(nbytes), 2) != 0:
            raise RuntimeError("hipMemcpy failed")
        audit = plan.audit()
        # algorithmic work: the far-field pairs (FP32 bin >= 1025, the audit's
        # count); the kernel also passes over the near-field pairs, whose
        # table entries are zero, and those are not counted
        far_pairs = audit["far_pairs"]
        flops = FLOPS_PER_PAIR_E * far_pairs
        achieved = flops / (far_avg * 1e-3) / 1e12
        hip.hipFree(dptr)
        plan.close()

        # the map set from host inputs without ASCII output: ag_maps_plain
        # (plan build, H2D, kernels, D2H into pinned host memory)
        hbuf = C.c_void_p()
        if hip.hipHostMalloc(C.byref(hbuf), C.c_size_t(nbytes), 0) != 0:
            raise RuntimeError("hipHostMalloc failed")
        try:
            host = np.ctypeslib.as_array(C.cast(hbuf, C.POINTER(C.c_float)),
                                         shape=(st.n_maps, st.n1[2], st.n1[1], st.n1[0]))
            ctx.maps_plain(st, out=host)
            walls = []
            for _ in range(max(3, args.steps // 4)):
                t1 = time.perf_counter()
                ctx.maps_plain(st, out=host)
                walls.append(time.perf_counter() - t1)
            host_equal = bool(np.array_equal(host.view(np.uint32), gpu_maps.view(np.uint32)))
        finally:
            hip.hipHostFree(hbuf)
        ctx.close()
        wall_excl = min(walls)

        # whole drop-in with ASCII output, for the wall time the reference reports
        cli_ok, wall_cli = False, None
        if not args.no_cli:
            os.chdir(work)
            try:
                t1 = time.perf_counter()
                p = subprocess.run([ag.CLI_PATH, "-p", info["gpf"], "-l", "gpu.glg"], capture_output=True)
                wall_cli = time.perf_counter() - t1
            finally:
                os.chdir(cwd)
            cli_ok = p.returncode == 0
    finally:
        shutil.rmtree(work, ignore_errors=True)

    large = None if args.no_large else large_leg()
    protein = None if args.no_protein else protein_leg()
    vina, vina_state = None, None
    if not args.no_vina and args.workload == "full":
        vina, vst, vmaps = vina_leg()
        vina_state = (vst, vmaps)
    batch = None if args.no_batch else batched_leg()
    analytic = None if args.no_analytic else analytic_leg(fit_steps=args.fit_steps)
    cpu = {"cpu_baseline": None, "accuracy": None}
    if not args.no_cpu_baseline and args.workload == "full":
        cpu = cpu_leg(st, gpu_maps, info["maps"], large=not args.no_large, vina=vina_state)
    value = ms_per_step / 1000.0
    # the whole map set's FP32 work over its kernels' time, when the CPU leg
    # has the exact count of pairs within 8 A: 16 FLOPs per point-atom pair
    # of the e pass, 16 + 4 per affinity map (+2 with a d map) per pair of
    # the affinity pass
    map_set_roofline = None
    acc = cpu.get("accuracy") or {}
    if acc.get("W_A") is not None:
        n_aff = sum(1 for k in range(st.n_maps) if k not in (st.elec_index, st.dsolv_index))
        f_a = 16 + 4 * n_aff + (2 if st.dsolv_index >= 0 else 0)
        ms_flops = 16.0 * P * N + f_a * acc["W_A"]
        t_k = sum(cell_ms) / len(cell_ms) + sum(near_ms) / len(near_ms) + far_avg
        ach = ms_flops / (t_k * 1e-3) / 1e12
        map_set_roofline = {"achieved": ach, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                            "frac": ach / FP32_PEAK_TFLOPS, "frac_nonpacked_peak": ach / FP32_PEAK_NONPACKED_TFLOPS,
                            "flops": ms_flops, "kernels_ms": t_k,
                            "flops_rule": f"16 W_E + F_A W_A, W_E = P*N, W_A = pairs within 8 A (oracle count), "
                                          f"F_A = 16 + 4*{n_aff} maps" + (" + 2 (d map)" if st.dsolv_index >= 0 else "")}
    out = {
        "metric": "map-set wall s (5k atoms,126^3,14 maps); batched & fwd+adj vs plain GPU",
        "value": value, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "value_is": "device-resident map-set time: ag_plan_execute on uploaded inputs (cell list, near "
                    "kernel, far kernel, status check), wall clock per step",
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{args.workload}: {N} atoms, {st.n1[0]}x{st.n1[1]}x{st.n1[2]} points, "
                               f"{st.n_maps} maps (13 affinity + e + d), spacing 0.375",
                   "atoms": N, "points": P, "maps": st.n_maps},
        "roofline": {"bound": "valu", "kernel": "k_far (electrostatics, all atoms beyond the near field)",
                     "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP32_PEAK_TFLOPS,
                     "frac_nonpacked_peak": achieved / FP32_PEAK_NONPACKED_TFLOPS,
                     "flops_per_launch": flops,
                     "flops_rule": "16 per far-field (point, atom) pair: the audit's far_pairs, FP32 bin >= 1025 "
                                   "(the near-field pairs the kernel passes over are not counted)",
                     "avg_launch_ms": far_avg,
                     "traffic": profiled_traffic() if args.workload == "full" else None,
                     "traffic_source": TRAFFIC_PROFILE if args.workload == "full" else None},
        "map_set_roofline": map_set_roofline,
        "kernels_ms": {"cell_list": sum(cell_ms) / len(cell_ms), "near": sum(near_ms) / len(near_ms),
                       "far": far_avg},
        "pair_evals_per_s": far_pairs / (far_avg * 1e-3),
        "pair_evals_per_s_is": "far-field (point, atom) pairs (the audit's far_pairs) over the far kernel's time",
        "map_set_pair_evals_per_s": P * N / (ms_per_step * 1e-3),
        "map_set_pair_evals_per_s_is": "all P*N (point, atom) pairs of the map set over its wall time per step",
        "bin_audit": audit,
        "frontend_s": frontend_s,
        "wall_excl_ascii_s": wall_excl,
        "wall_excl_ascii_is": "ag_maps_plain from host inputs: plan build, H2D, kernels, D2H into pinned memory "
                              "(front end not included, see frontend_s)",
        "wall_excl_ascii_equals_device_run": host_equal,
        "wall_with_ascii_output_s": wall_cli if cli_ok else None,
        "wall_with_ascii_output_is": "the autogrid4 drop-in end to end: GPF/PDBQT read, tables, device, "
                                     "15 ASCII maps, .fld, .xyz, log",
        "accuracy": cpu["accuracy"],
        "large": large,
        "protein": protein,
        "vina": vina,
        "batched": batch,
        "analytic": analytic,
        "cpu_baseline": cpu["cpu_baseline"],
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
