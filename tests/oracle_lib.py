"""ctypes access to the FP64 oracle (oracle/liboracle_ag.so) for the tests.

The oracle is the checker: tests compare the engine's device output with it
and pin it against the reference autogrid4 binary and the reference's own
fixtures.  It is never part of the engine.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

import autogrid_amd as ag
from autogrid_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.environ.get("AG_ORACLE_SO") or os.path.join(ORACLE_DIR, "liboracle_ag.so")
REF_BIN = os.path.join(ORACLE_DIR, "_ref", "autogrid4")

_lib = None


def olib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            subprocess.run(["make", "-C", ORACLE_DIR, "-s"], check=True)
        L = C.CDLL(ORACLE_SO)
        vp = C.c_void_p
        L.orc_vdw_tables.argtypes = [C.POINTER(_abi.Setup), C.POINTER(_abi.MapSet), vp]
        L.orc_sol_fn.argtypes = [C.c_double, vp]
        L.orc_sol_fn.restype = None
        L.orc_r_eps.argtypes = [vp]
        L.orc_r_eps.restype = None
        L.orc_ddd.argtypes = [C.c_double]
        L.orc_ddd.restype = C.c_double
        L.orc_bond_vectors.argtypes = [C.POINTER(_abi.Setup), C.c_int, vp, vp, vp, vp, vp, vp, vp]
        L.orc_maps.argtypes = [C.POINTER(_abi.Receptor), C.POINTER(_abi.Grid), C.POINTER(_abi.MapSet),
                               C.POINTER(_abi.Tables), vp, vp, vp, C.c_int, C.c_int]
        L.orc_xs_radius.argtypes = [C.c_char_p, C.POINTER(C.c_double)]
        L.orc_vina_tables.argtypes = [C.POINTER(_abi.Setup), C.POINTER(_abi.MapSet), vp, vp, vp]
        L.orc_maps_vina.argtypes = L.orc_maps.argtypes
        L.orc_analytic_maps.argtypes = [C.POINTER(_abi.Receptor), C.POINTER(_abi.Grid), C.POINTER(_abi.MapSet),
                                        vp, C.c_int, C.c_double, C.c_double, C.c_int, vp, vp]
        L.orc_count_pairs.argtypes = [C.POINTER(_abi.Receptor), C.POINTER(_abi.Grid), C.c_int, C.c_int]
        L.orc_count_pairs.restype = C.c_longlong
        L.orc_format_value.argtypes = [C.c_double, C.c_char_p]
        L.orc_format_rmin.argtypes = [C.c_double, C.c_char_p]
        _lib = L
    return _lib


def tables(st: ag.PreparedState):
    """The oracle's (e_vdw_hb, sol_fn, r_eps) for a parsed state."""
    rec, grid, ms, tab, setup = st.structs()
    T, M = st.e_vdw_hb.shape[0], st.n_atom_maps
    ev = np.zeros((T, M, 1024), np.float32)
    rc = olib().orc_vdw_tables(C.byref(setup), C.byref(ms), ev.ctypes.data)
    if rc != 0:
        raise RuntimeError("oracle table build stopped")
    sol = np.zeros(8192)
    olib().orc_sol_fn(st.coeff_desolv, sol.ctypes.data)
    reps = np.zeros(8192)
    if st.dddiel:
        olib().orc_r_eps(reps.ctypes.data)
    return ev, sol, reps


def bond_vectors(st: ag.PreparedState):
    rec, grid, ms, tab, setup = st.structs()
    n = st.n_atoms
    rexp = np.zeros(n, np.int32)
    rv = np.zeros((n, 3))
    rv2 = np.zeros((n, 3))
    dis = np.zeros(n, np.uint8)
    coord = np.ascontiguousarray(st.coord)
    typ = np.ascontiguousarray(st.type, np.int32)
    hb = np.ascontiguousarray(st.hbond, np.int32)
    rc = olib().orc_bond_vectors(C.byref(setup), n, coord.ctypes.data, typ.ctypes.data, hb.ctypes.data,
                                 rexp.ctypes.data, rv.ctypes.data, rv2.ctypes.data, dis.ctypes.data)
    if rc != 0:
        raise RuntimeError("oracle bond-vector pass stopped")
    return rexp, rv, rv2, dis


def oracle_state(st: ag.PreparedState) -> ag.PreparedState:
    """A copy of st whose tables and bond vectors come from the oracle."""
    import dataclasses
    ev, sol, reps = tables(st)
    rexp, rv, rv2, dis = bond_vectors(st)
    return dataclasses.replace(st, e_vdw_hb=ev, sol_fn=sol, r_eps=reps, rexp=rexp, rvector=rv, rvector2=rv2,
                               disorder=dis, _keep=[])


def count_pairs(st: ag.PreparedState, z0: int = 0, z1: int | None = None) -> int:
    """W_A over planes z0..z1: (point, atom) pairs with r <= 8 A (Loop 2)."""
    rec, grid, _, _, _ = st.structs()
    n = olib().orc_count_pairs(C.byref(rec), C.byref(grid), z0, st.n1[2] if z1 is None else z1)
    if n < 0:
        raise RuntimeError("oracle: bad plane range")
    return int(n)


def maps(st: ag.PreparedState, z0: int = 0, z1: int | None = None, fmap=False, cmap=False):
    """FP64 maps [n_maps, n1z, n1y, n1x] (planes z0..z1) of the state."""
    rec, grid, ms, tab, setup = st.structs()
    nz, ny, nx = st.n1[2], st.n1[1], st.n1[0]
    z1 = nz if z1 is None else z1
    out = np.zeros((st.n_maps, nz, ny, nx))
    fm = np.zeros((nz, ny, nx)) if fmap else None
    cm = np.zeros((nz, ny, nx), np.int32) if cmap else None
    rc = olib().orc_maps(C.byref(rec), C.byref(grid), C.byref(ms), C.byref(tab), out.ctypes.data,
                         fm.ctypes.data if fmap else None, cm.ctypes.data if cmap else None, z0, z1)
    if rc != 0:
        raise RuntimeError("oracle: no closestH atom was found")
    return (out, fm, cm) if (fmap or cmap) else out


def format_value(e: float) -> bytes:
    buf = C.create_string_buffer(64)
    n = olib().orc_format_value(float(e), buf)
    return buf.raw[:n]


def format_map(values) -> bytes:
    return b"".join(format_value(float(v)) for v in np.asarray(values).reshape(-1))


def format_rmin_map(values) -> bytes:
    buf = C.create_string_buffer(64)
    out = []
    for v in np.asarray(values).reshape(-1):
        n = olib().orc_format_rmin(float(v), buf)
        out.append(buf.raw[:n])
    return b"".join(out)


def format_count_map(values) -> bytes:
    return b"".join(b"%d\n" % int(v) for v in np.asarray(values).reshape(-1))


def xs_radius(type_name: str):
    r = C.c_double()
    found = olib().orc_xs_radius(type_name.encode(), C.byref(r))
    return float(r.value), bool(found)


def vina_tables(st: ag.PreparedState, lig_types, rec_types) -> np.ndarray:
    """The oracle's Vina tables [T, M, 1024] for a state's pairs and type names."""
    _, _, ms, _, setup = st.structs()
    xs_map = np.array([xs_radius(t)[0] for t in lig_types], np.float64)
    xs_rec = np.array([xs_radius(t)[0] for t in rec_types], np.float64)
    out = np.zeros((len(rec_types), st.n_atom_maps, 1024), np.float32)
    olib().orc_vina_tables(C.byref(setup), C.byref(ms), xs_map.ctypes.data, xs_rec.ctypes.data, out.ctypes.data)
    return out


def maps_vina(st: ag.PreparedState, z0: int = 0, z1: int | None = None, fmap=False, cmap=False):
    """FP64 Vina map set [n_maps, n1z, n1y, n1x] (planes z0..z1) of a state holding Vina tables."""
    rec, grid, ms, tab, setup = st.structs()
    nz, ny, nx = st.n1[2], st.n1[1], st.n1[0]
    z1 = nz if z1 is None else z1
    out = np.zeros((st.n_maps, nz, ny, nx))
    fm = np.zeros((nz, ny, nx)) if fmap else None
    cm = np.zeros((nz, ny, nx), np.int32) if cmap else None
    olib().orc_maps_vina(C.byref(rec), C.byref(grid), C.byref(ms), C.byref(tab), out.ctypes.data,
                         fm.ctypes.data if fmap else None, cm.ctypes.data if cmap else None, z0, z1)
    return (out, fm, cm) if (fmap or cmap) else out


def gpf_types(gpf_path: str):
    """(ligand_types, receptor_types) as written in a GPF."""
    lig, recp = [], []
    for line in open(gpf_path):
        w = line.split("#")[0].split()
        if w and w[0] == "ligand_types":
            lig = w[1:]
        elif w and w[0] == "receptor_types":
            recp = w[1:]
    return lig, recp


def analytic_maps(st: ag.PreparedState, points, rij=None, eps0=None, weights=None, coord=None, charge=None):
    """FP64 analytic model at the given flat point indices: [n_maps, npts].

    rij / eps0 ([M, T], eps0 unweighted) and weights (vdw, hbond, estat,
    desolv) default to the state's; bond vectors are rebuilt by the oracle
    from the (possibly perturbed) coordinates."""
    import dataclasses
    w = np.array(st.weights if weights is None else weights, np.float64)
    pairs = st.pairs.copy()
    wk = np.where(pairs["hbonder"] > 0, st.weights[1], st.weights[0])
    e0 = pairs["eps"] / wk if eps0 is None else np.asarray(eps0, np.float64)
    if rij is not None:
        pairs["Rij"] = rij
    pairs["eps"] = e0 * np.where(pairs["hbonder"] > 0, w[1], w[0])
    s2 = dataclasses.replace(st, coord=st.coord if coord is None else np.asarray(coord, np.float64),
                             charge=st.charge if charge is None else np.asarray(charge, np.float64),
                             coeff_estat=float(w[2]), _keep=[])
    rexp, rv, rv2, dis = bond_vectors(s2)
    s2 = dataclasses.replace(s2, rexp=rexp, rvector=rv, rvector2=rv2, disorder=dis, _keep=[])
    rec, grid, ms, tab, setup = s2.structs()
    M, T = pairs.shape
    pa = (_abi.Pair * max(1, M * T))()
    for i, p in enumerate(pairs.reshape(-1)):
        pa[i] = _abi.Pair(*[p[k].item() for k in ("cA", "cB", "Rij", "eps", "xA", "xB", "hbonder")], 0)
    pts = np.ascontiguousarray(points, np.int64)
    out = np.zeros((st.n_maps, pts.size))
    rc = olib().orc_analytic_maps(C.byref(rec), C.byref(grid), C.byref(ms), C.cast(pa, C.c_void_p), T,
                                  s2.r_smooth, float(w[3]), pts.size, pts.ctypes.data, out.ctypes.data)
    if rc != 0:
        raise RuntimeError("oracle analytic model refused the state")
    return out
