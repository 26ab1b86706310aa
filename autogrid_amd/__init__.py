"""MI355X-native AutoGrid4 grid-map engine.

Host mirror of the autogrid4 interface over the C-ABI of
``libautogrid_amd.so``:

* :func:`run_gpf` -- the whole drop-in (``autogrid4 -p x.gpf -l x.glg``):
  GPF/PDBQT in, ``.map``/``.fld``/``.xyz`` out (main.cpp:214-3582).
* :class:`Job` -- the front end: GPF keyword semantics (main.cpp:778-1819),
  tables and bond vectors (main.cpp:1607-1639, 2040-2736) as numpy arrays.
* :class:`Context` -- one GPU; :meth:`Context.maps_plain` runs the per-point
  map computation (main.cpp:2805-3431) on the device.

Everything numeric runs in the native library; this module only moves
arrays.  It raises when the library is missing -- there is no fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

from . import _abi
from ._abi import AgError, check, lib

__all__ = ["AgError", "Context", "Job", "Plan", "PreparedState", "run_gpf", "format_map_values", "xs_radius",
           "build_tables", "bond_vectors", "bond_vectors_batch", "Analytic",
           "vina_tables", "AG_POTENTIAL_AD4", "AG_POTENTIAL_VINA",
           "LIB_PATH", "CLI_PATH"]

LIB_PATH = _abi.LIB_PATH
AG_POTENTIAL_AD4 = _abi.AG_POTENTIAL_AD4
AG_POTENTIAL_VINA = _abi.AG_POTENTIAL_VINA
CLI_PATH = _abi.CLI_PATH


@dataclass
class PreparedState:
    """The seam state of include/autogrid_amd.h as numpy arrays."""
    coord: np.ndarray            # [n,3] float64, grid frame
    charge: np.ndarray           # [n] float64
    vol: np.ndarray
    solpar: np.ndarray
    type: np.ndarray             # [n] int32
    hbond: np.ndarray            # [n] int32
    rexp: np.ndarray             # [n] int32
    rvector: np.ndarray          # [n,3] float64
    rvector2: np.ndarray
    disorder: np.ndarray         # [n] uint8
    n1: tuple
    ne: tuple
    spacing: float
    maps: np.ndarray             # structured: is_covalent,is_hbonder,hbond,solpar_probe,vol_probe
    elec_index: int
    dsolv_index: int
    n_maps: int
    separate_desolvation: int
    disorder_h: int
    hydrogen_type: int
    dddiel: int
    inv_diel_cal: float
    coeff_estat: float
    solpar_q: float
    e_vdw_hb: np.ndarray         # [T, M, 1024] float32
    sol_fn: np.ndarray           # [8192] float64
    r_eps: np.ndarray            # [8192] float64
    cov_pos: tuple = (0.0, 0.0, 0.0)
    cov_half_width: float = 1.0
    cov_barrier: float = 1000.0
    floating_grid: int = 0
    constriction_grid: int = 0
    constriction_cutoff: float = 7.5
    bhtree: int = 0              # ag_mapset.bhtree: 0 default build, 1 USE_BHTREE, 2 USE_BHTREE + interior
    # ag_setup fields (what the table / bond-vector builders read)
    pairs: np.ndarray | None = None
    r_smooth: float = 0.5
    coeff_desolv: float = 0.1322
    t_types: dict = field(default_factory=dict)
    weights: tuple = (0.1662, 0.1209, 0.1406, 0.1322)   # library coeff_vdW, hbond, estat, desolv
    _keep: list = field(default_factory=list, repr=False)

    @property
    def n_atoms(self) -> int:
        return int(self.coord.shape[0])

    @property
    def points(self) -> int:
        return int(self.n1[0]) * int(self.n1[1]) * int(self.n1[2])

    @property
    def n_atom_maps(self) -> int:
        return int(len(self.maps))

    def receptor_grid(self):
        """ctypes (Receptor, Grid) pointing into this state, and the arrays
        that must outlive them."""
        arrs = dict(
            coord=np.ascontiguousarray(self.coord, np.float64), charge=np.ascontiguousarray(self.charge, np.float64),
            vol=np.ascontiguousarray(self.vol, np.float64), solpar=np.ascontiguousarray(self.solpar, np.float64),
            type=np.ascontiguousarray(self.type, np.int32), hbond=np.ascontiguousarray(self.hbond, np.int32),
            rexp=np.ascontiguousarray(self.rexp, np.int32), rvector=np.ascontiguousarray(self.rvector, np.float64),
            rvector2=np.ascontiguousarray(self.rvector2, np.float64),
            disorder=np.ascontiguousarray(self.disorder, np.uint8))
        P = _abi._ptr
        rec = _abi.Receptor(self.n_atoms, P(arrs["coord"], C.c_double), P(arrs["charge"], C.c_double),
                            P(arrs["vol"], C.c_double), P(arrs["solpar"], C.c_double),
                            P(arrs["type"], C.c_int32), P(arrs["hbond"], C.c_int32),
                            P(arrs["rexp"], C.c_int32), P(arrs["rvector"], C.c_double),
                            P(arrs["rvector2"], C.c_double), P(arrs["disorder"], C.c_uint8))
        grid = _abi.Grid((C.c_int32 * 3)(*self.n1), (C.c_int32 * 3)(*self.ne), float(self.spacing))
        return rec, grid, arrs

    def structs(self):
        """ctypes (Receptor, Grid, MapSet, Tables, Setup) pointing into this state."""
        rec, grid, arrs = self.receptor_grid()
        arrs.update(e_vdw_hb=np.ascontiguousarray(self.e_vdw_hb, np.float32),
                    sol_fn=np.ascontiguousarray(self.sol_fn, np.float64),
                    r_eps=np.ascontiguousarray(self.r_eps, np.float64))
        P = _abi._ptr
        md = (_abi.MapDesc * max(1, self.n_atom_maps))()
        for i, m in enumerate(self.maps):
            md[i] = _abi.MapDesc(int(m["is_covalent"]), int(m["is_hbonder"]), int(m["hbond"]), 0,
                                 float(m["solpar_probe"]), float(m["vol_probe"]))
        ms = _abi.MapSet(self.n_atom_maps, C.cast(md, C.POINTER(_abi.MapDesc)), self.elec_index, self.dsolv_index,
                         self.n_maps, self.separate_desolvation, self.disorder_h, self.hydrogen_type,
                         self.dddiel, self.floating_grid, self.constriction_grid, int(self.bhtree),
                         self.inv_diel_cal,
                         self.coeff_estat, self.solpar_q, (C.c_double * 3)(*self.cov_pos),
                         self.cov_half_width, self.cov_barrier, self.constriction_cutoff)
        tab = _abi.Tables(int(self.e_vdw_hb.shape[0]), 0, P(arrs["e_vdw_hb"], C.c_float),
                          P(arrs["sol_fn"], C.c_double), P(arrs["r_eps"], C.c_double))
        setup = _abi.Setup()
        setup.n_rec_types = int(self.e_vdw_hb.shape[0])
        setup.n_atom_maps = self.n_atom_maps
        keep = [arrs, md]
        if self.pairs is not None:
            pa = (_abi.Pair * max(1, self.pairs.size))()
            view = np.frombuffer(pa, dtype=PAIR_ABI_DTYPE)
            flat = self.pairs.reshape(-1)
            for k in ("cA", "cB", "Rij", "eps", "xA", "xB", "hbonder"):
                view[k][:flat.size] = flat[k]
            setup.pairs = C.cast(pa, C.POINTER(_abi.Pair))
            keep.append(pa)
        setup.r_smooth = self.r_smooth
        setup.coeff_desolv = self.coeff_desolv
        setup.dddiel = self.dddiel
        setup.disorder_h = self.disorder_h
        for k in ("t_hydrogen", "t_nonhb_hydrogen", "t_carbon", "t_arom_carbon", "t_oxygen", "t_sulphur",
                  "t_nitrogen", "t_nonhb_nitrogen"):
            setattr(setup, k, int(self.t_types.get(k, -1)))
        self._keep[:] = keep   # the structs point into these arrays
        return rec, grid, ms, tab, setup


MAP_DTYPE = np.dtype([("is_covalent", np.int32), ("is_hbonder", np.int32), ("hbond", np.int32),
                      ("solpar_probe", np.float64), ("vol_probe", np.float64)])
PAIR_DTYPE = np.dtype([("cA", np.float64), ("cB", np.float64), ("Rij", np.float64), ("eps", np.float64),
                       ("xA", np.int32), ("xB", np.int32), ("hbonder", np.int32)])
# the C layout of ag_pair (one int32 of padding)
PAIR_ABI_DTYPE = np.dtype([("cA", np.float64), ("cB", np.float64), ("Rij", np.float64), ("eps", np.float64),
                           ("xA", np.int32), ("xB", np.int32), ("hbonder", np.int32), ("reserved", np.int32)])


def read_binary_map(path: str, n1) -> np.ndarray:
    """A map of the binary side-output: raw little-endian float32, x fastest,
    as [n1z, n1y, n1x]."""
    return np.fromfile(path, dtype="<f4").reshape(n1[2], n1[1], n1[0])


def _arr(ptr, n, dtype, shape=None):
    if n == 0 or not ptr:
        return np.zeros(shape if shape else (0,), dtype)
    a = np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)
    return a.reshape(shape) if shape else a


class Job:
    """A GPF read by the engine's front end (ag_job_open)."""

    def __init__(self, gpf_path: str, glg_path: str | None = None, program: str = "autogrid4",
                 files: bool = False, bhtree: bool = False, binary_maps: bool = False):
        h = C.c_void_p()
        flags = ((0 if files else _abi.AG_JOB_NO_FILES) | (_abi.AG_JOB_BHTREE if bhtree else 0)
                 | (_abi.AG_JOB_BINARY_MAPS if binary_maps else 0))
        check(lib().ag_job_open(os.fsencode(gpf_path), os.fsencode(glg_path) if glg_path else None,
                                program.encode(), flags, C.byref(h)))
        self._h = h
        self.prepared = False

    def close(self):
        if getattr(self, "_h", None):
            lib().ag_job_close(self._h)
            self._h = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def prepare(self) -> "Job":
        check(lib().ag_job_prepare(self._h))
        self.prepared = True
        return self

    def prepare_vina(self) -> "Job":
        """prepare(), with the job's tables replaced by Vina-potential tables."""
        check(lib().ag_job_prepare_vina(self._h))
        self.prepared = True
        return self

    @property
    def points(self) -> int:
        return int(lib().ag_job_points(self._h))

    def write_maps(self, maps: np.ndarray, fmap: np.ndarray | None = None, cmap: np.ndarray | None = None,
                   threads: int = 4) -> None:
        """Write the job's output files from host maps (ag_job_write_maps)."""
        maps = np.ascontiguousarray(maps, np.float32)
        fm = None if fmap is None else np.ascontiguousarray(fmap, np.float32)
        cm = None if cmap is None else np.ascontiguousarray(cmap, np.int32)
        check(lib().ag_job_write_maps(self._h, maps.ctypes.data, fm.ctypes.data if fm is not None else None,
                                      cm.ctypes.data if cm is not None else None, int(threads)))

    def state(self) -> PreparedState:
        """Copy of the parsed (and, after prepare(), prepared) state."""
        rec, grid, ms, tab, setup = (_abi.Receptor(), _abi.Grid(), _abi.MapSet(), _abi.Tables(), _abi.Setup())
        check(lib().ag_job_views(self._h, C.byref(rec), C.byref(grid), C.byref(ms), C.byref(tab), C.byref(setup)))
        n = rec.n_atoms
        T, M = tab.n_rec_types, ms.n_atom_maps
        maps = np.zeros(M, MAP_DTYPE)
        if M and ms.maps:
            for i in range(M):
                d = ms.maps[i]
                maps[i] = (d.is_covalent, d.is_hbonder, d.hbond, d.solpar_probe, d.vol_probe)
        pairs = None
        if setup.pairs and T * M:
            pairs = np.zeros((M, T), PAIR_DTYPE)
            for i in range(M * T):
                p = setup.pairs[i]
                pairs.reshape(-1)[i] = (p.cA, p.cB, p.Rij, p.eps, p.xA, p.xB, p.hbonder)
        w = _abi.Weights()
        check(lib().ag_job_weights(self._h, C.byref(w)))
        return PreparedState(
            weights=(w.vdw, w.hbond, w.estat, w.desolv),
            coord=_arr(rec.coord, 3 * n, np.float64, (n, 3)), charge=_arr(rec.charge, n, np.float64),
            vol=_arr(rec.vol, n, np.float64), solpar=_arr(rec.solpar, n, np.float64),
            type=_arr(rec.type, n, np.int32), hbond=_arr(rec.hbond, n, np.int32),
            rexp=_arr(rec.rexp, n, np.int32), rvector=_arr(rec.rvector, 3 * n, np.float64, (n, 3)),
            rvector2=_arr(rec.rvector2, 3 * n, np.float64, (n, 3)), disorder=_arr(rec.disorder, n, np.uint8),
            n1=tuple(grid.n1), ne=tuple(grid.ne), spacing=grid.spacing, maps=maps,
            elec_index=ms.elec_index, dsolv_index=ms.dsolv_index, n_maps=ms.n_maps,
            separate_desolvation=ms.separate_desolvation, disorder_h=ms.disorder_h,
            hydrogen_type=ms.hydrogen_type, dddiel=ms.dddiel, inv_diel_cal=ms.inv_diel_cal,
            coeff_estat=ms.coeff_estat, solpar_q=ms.solpar_q,
            e_vdw_hb=_arr(tab.e_vdw_hb, T * M * _abi.AG_NEINT, np.float32, (T, M, _abi.AG_NEINT)),
            sol_fn=_arr(tab.sol_fn, _abi.AG_NDIEL, np.float64), r_eps=_arr(tab.r_eps, _abi.AG_NDIEL, np.float64),
            cov_pos=tuple(ms.cov_pos), cov_half_width=ms.cov_half_width, cov_barrier=ms.cov_barrier,
            floating_grid=ms.floating_grid, constriction_grid=ms.constriction_grid,
            constriction_cutoff=ms.constriction_cutoff, bhtree=ms.bhtree, pairs=pairs, r_smooth=setup.r_smooth,
            coeff_desolv=setup.coeff_desolv,
            t_types={k: getattr(setup, k) for k in ("t_hydrogen", "t_nonhb_hydrogen", "t_carbon", "t_arom_carbon",
                                                    "t_oxygen", "t_sulphur", "t_nitrogen", "t_nonhb_nitrogen")})


class Context:
    """One GPU and its stream (ag_context_create)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(lib().ag_context_create(C.byref(h), int(device)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().ag_context_destroy(self._h)
            self._h = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    def maps_plain(self, st: PreparedState, out: np.ndarray | None = None,
                   potential: int = _abi.AG_POTENTIAL_AD4) -> np.ndarray:
        """[n_maps, n1z, n1y, n1x] float32 maps for the state, computed on the GPU."""
        rec, grid, ms, tab, _ = st.structs()
        entry = lib().ag_maps_vina if potential == _abi.AG_POTENTIAL_VINA else lib().ag_maps_plain
        nz, ny, nx = st.n1[2], st.n1[1], st.n1[0]
        if out is None:
            out = np.empty((st.n_maps, nz, ny, nx), np.float32)
        assert out.dtype == np.float32 and out.flags.c_contiguous and out.size == st.n_maps * st.points
        fmap = np.empty((nz, ny, nx), np.float32) if st.floating_grid else None
        cmap = np.empty((nz, ny, nx), np.int32) if st.constriction_grid else None
        check(entry(self._h, C.byref(rec), C.byref(grid), C.byref(ms), C.byref(tab),
                    out.ctypes.data, fmap.ctypes.data if fmap is not None else None,
                    cmap.ctypes.data if cmap is not None else None, 0))
        self.last_fmap, self.last_cmap = fmap, cmap
        return out

    def maps_vina(self, st: PreparedState, out: np.ndarray | None = None) -> np.ndarray:
        """Vina-potential map set (ag_maps_vina); st.e_vdw_hb must hold Vina tables."""
        return self.maps_plain(st, out, potential=_abi.AG_POTENTIAL_VINA)

    def maps_plain_device(self, st: PreparedState, out_ptr: int) -> None:
        """Same, writing into device memory at out_ptr ([n_maps][P] float32)."""
        rec, grid, ms, tab, _ = st.structs()
        check(lib().ag_maps_plain(self._h, C.byref(rec), C.byref(grid), C.byref(ms), C.byref(tab),
                                  C.c_void_p(out_ptr), None, None, 1))

    def maps_plain_noalloc(self, st: PreparedState, out_ptr: int, fmap_ptr: int | None = None,
                           cmap_ptr: int | None = None) -> None:
        """The map set into device memory at out_ptr ([n_maps][P] float32)
        with no device allocation (ag_maps_plain_noalloc): every buffer
        comes from the device's fixed scratch array."""
        rec, grid, ms, tab, _ = st.structs()
        check(lib().ag_maps_plain_noalloc(self._h, C.byref(rec), C.byref(grid), C.byref(ms), C.byref(tab),
                                          C.c_void_p(out_ptr), C.c_void_p(fmap_ptr), C.c_void_p(cmap_ptr)))

    def batch(self, states: list, tune: dict | None = None, potential: int = _abi.AG_POTENTIAL_AD4) -> "Plan":
        """Upload many conformations sharing states[0]'s map set and tables
        (ag_batch_create); tune None runs the autotuner."""
        return Plan(self, states, potential, batch=True, tune=tune)

    def maps_batch(self, states: list, tune: dict | None = None,
                   potential: int = _abi.AG_POTENTIAL_AD4, outs: list | None = None,
                   fmaps: list | None = None, cmaps: list | None = None) -> list:
        """Maps of many conformations (ag_maps_batch): a list of
        [n_maps, n1z, n1y, n1x] float32 arrays.  Caller-supplied outputs
        (outs, and fmaps / cmaps when the map set asks for a floating or
        constriction grid) must be C-contiguous arrays of the conformations'
        shapes, float32 (int32 for cmaps); when all of them are pinned host
        memory the library overlaps the copies with the computation."""
        st0 = states[0]
        recs, grids, ms, tab, keep = _batch_structs(states)
        shapes = [(st.n1[2], st.n1[1], st.n1[0]) for st in states]
        outs, fm, cm, ptrs, fptrs, cptrs = _host_outputs(shapes, st0.n_maps, bool(st0.floating_grid),
                                                         bool(st0.constriction_grid), outs, fmaps, cmaps)
        t = _tune_struct(tune)
        check(lib().ag_maps_batch(self._h, len(states), recs, grids, C.byref(ms), C.byref(tab), ptrs, fptrs, cptrs,
                                  int(potential), C.byref(t) if t is not None else None))
        self.last_fmaps, self.last_cmaps = fm, cm
        return outs

    def plan(self, st: PreparedState, potential: int = _abi.AG_POTENTIAL_AD4) -> "Plan":
        """Upload the state to this context's device (ag_plan_create_potential)."""
        return Plan(self, st, potential)

    def timing(self) -> dict:
        """HIP-event milliseconds of the last plan execution on this context."""
        ms = (C.c_float * 4)()
        check(lib().ag_context_timing(self._h, ms))
        return {"cell_ms": ms[0], "near_ms": ms[1], "far_ms": ms[2], "total_ms": ms[3]}

    def run_gpf(self, gpf_path: str, glg_path: str | None = None, program: str = "autogrid4",
                bhtree: bool = False, binary_maps: bool = False) -> None:
        flags = (_abi.AG_JOB_BHTREE if bhtree else 0) | (_abi.AG_JOB_BINARY_MAPS if binary_maps else 0)
        check(lib().ag_run_gpf_flags(self._h, os.fsencode(gpf_path), os.fsencode(glg_path) if glg_path else None,
                                     program.encode(), flags))


def _host_outputs(shapes: list, n_maps: int, want_fmap: bool, want_cmap: bool, outs, fmaps, cmaps):
    """Host output arrays of a batch (allocated where not given, checked
    where given: C-contiguous, float32 maps [n_maps, n1z, n1y, n1x] and
    fmaps [n1z, n1y, n1x], int32 cmaps) and ctypes pointer arrays to them."""
    def outputs(given, wanted, dtype, lead):
        if not wanted:
            return None
        if given is None:
            return [np.empty(lead + sh, dtype) for sh in shapes]
        if len(given) != len(shapes):
            raise ValueError(f"{len(given)} output arrays for {len(shapes)} conformations")
        for i, (a, sh) in enumerate(zip(given, shapes)):
            want = lead + tuple(sh)
            if not (isinstance(a, np.ndarray) and a.dtype == dtype and a.flags.c_contiguous and a.shape == want):
                raise ValueError(f"output {i} must be a C-contiguous {np.dtype(dtype).name} array of "
                                 f"shape {want}")
        return list(given)
    outs = outputs(outs, True, np.float32, (n_maps,))
    fm = outputs(fmaps, want_fmap, np.float32, ())
    cm = outputs(cmaps, want_cmap, np.int32, ())
    n = len(shapes)
    ptrs = (C.c_void_p * n)(*[o.ctypes.data for o in outs])
    fptrs = (C.c_void_p * n)(*[a.ctypes.data for a in fm]) if fm else None
    cptrs = (C.c_void_p * n)(*[a.ctypes.data for a in cm]) if cm else None
    return outs, fm, cm, ptrs, fptrs, cptrs


def _tune_struct(tune: dict | None):
    if tune is None:
        return None
    return _abi.Tune(float(tune.get("cell_size", 4.0)), int(bool(tune.get("cube", False))), 0, 0.0)


_RECEPTOR_FIELDS = (("coord", np.float64), ("charge", np.float64), ("vol", np.float64), ("solpar", np.float64),
                    ("type", np.int32), ("hbond", np.int32), ("rexp", np.int32), ("rvector", np.float64),
                    ("rvector2", np.float64), ("disorder", np.uint8))


def _batch_structs(states: list):
    """ctypes arrays of Receptor and Grid for a list of states, the first
    state's MapSet and Tables, and the list that keeps their arrays alive.
    The receptors are filled with plain addresses (ReceptorAddr, the same
    layout), one integer per field instead of a typed-pointer cast."""
    n = len(states)
    recs = (_abi.ReceptorAddr * n)()
    grids = (_abi.Grid * n)()
    _, _, ms, tab, _ = states[0].structs()
    keep = [recs, list(states[0]._keep)]
    for i, st in enumerate(states):
        r = recs[i]
        r.n_atoms = st.n_atoms
        arrs = []
        for name, dt in _RECEPTOR_FIELDS:
            a = np.ascontiguousarray(getattr(st, name), dt)
            arrs.append(a)
            setattr(r, name, a.ctypes.data)
        keep.append(arrs)
        g = grids[i]
        g.n1[:] = st.n1
        g.ne[:] = st.ne
        g.spacing = float(st.spacing)
    return C.cast(recs, C.POINTER(_abi.Receptor)), grids, ms, tab, keep


class Plan:
    """Receptor conformation(s), grid(s), map set and tables resident on one device."""

    def __init__(self, ctx: Context, st, potential: int = _abi.AG_POTENTIAL_AD4, batch: bool = False,
                 tune: dict | None = None):
        h = C.c_void_p()
        if batch:
            states = list(st)
            recs, grids, ms, tab, keep = _batch_structs(states)
            t = _tune_struct(tune)
            check(lib().ag_batch_create(ctx.handle, len(states), recs, grids, C.byref(ms), C.byref(tab),
                                        int(potential), C.byref(t) if t is not None else None, C.byref(h)))
            self.n_maps = states[0].n_maps
        else:
            rec, grid, ms, tab, _ = st.structs()
            check(lib().ag_plan_create_potential(ctx.handle, C.byref(rec), C.byref(grid), C.byref(ms),
                                                 C.byref(tab), int(potential), C.byref(h)))
            self.n_maps = st.n_maps
            states = [st]
        self.shapes = [(s.n1[2], s.n1[1], s.n1[0]) for s in states]
        self.want_fmap = bool(states[0].floating_grid)
        self.want_cmap = bool(states[0].constriction_grid)
        self._h = h
        self.ctx = ctx
        n = lib().ag_plan_conformations(h)
        self.map_offsets = np.zeros(n, np.int64)
        self.point_offsets = np.zeros(n, np.int64)
        mf, pts = C.c_int64(), C.c_int64()
        check(lib().ag_plan_layout(h, self.map_offsets.ctypes.data, self.point_offsets.ctypes.data,
                                   C.byref(mf), C.byref(pts)))
        self.map_floats = int(mf.value)
        self.points = int(pts.value)

    def tune(self) -> dict:
        """The near-kernel layout in use and the autotuner's time (ag_plan_tune)."""
        t = _abi.Tune()
        check(lib().ag_plan_tune(self._h, C.byref(t)))
        return {"cell_size": t.cell_size, "cube": bool(t.cube), "tune_ms": t.tune_ms}

    def execute(self, maps_dev_ptr: int, fmap_dev_ptr: int | None = None, cmap_dev_ptr: int | None = None) -> None:
        """All maps into device memory at maps_dev_ptr ([n_maps][P] float32)."""
        check(lib().ag_plan_execute(self._h, C.c_void_p(maps_dev_ptr), C.c_void_p(fmap_dev_ptr),
                                    C.c_void_p(cmap_dev_ptr)))

    def execute_host(self, maps_dev_ptr: int, outs: list | None = None, fmaps: list | None = None,
                     cmaps: list | None = None, fmap_dev_ptr: int | None = None,
                     cmap_dev_ptr: int | None = None) -> list:
        """Execute into the device buffers and copy every conformation's
        maps (and fmap / cmap) into host arrays (ag_plan_execute_host):
        the maps other than e are copied while the far-field kernel runs.
        Returns the list of [n_maps, n1z, n1y, n1x] float32 arrays; the
        fmap / cmap arrays are kept in last_fmaps / last_cmaps."""
        outs, fm, cm, ptrs, fptrs, cptrs = _host_outputs(self.shapes, self.n_maps, self.want_fmap,
                                                         self.want_cmap, outs, fmaps, cmaps)
        check(lib().ag_plan_execute_host(self._h, C.c_void_p(maps_dev_ptr), C.c_void_p(fmap_dev_ptr),
                                         C.c_void_p(cmap_dev_ptr), ptrs, fptrs, cptrs))
        self.last_fmaps, self.last_cmaps = fm, cm
        return outs

    def audit(self) -> dict:
        """Table-bin audit against the reference's FP64 distances (ag_plan_audit)."""
        v = (C.c_int64 * 4)()
        check(lib().ag_plan_audit(self._h, C.byref(v)))
        return {"far_pairs": v[0], "far_bin_flips": v[1], "near_unguarded_flips": v[2],
                "max_flips_per_point": v[3]}

    def close(self):
        if getattr(self, "_h", None):
            lib().ag_plan_destroy(self._h)
            self._h = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def run_gpf(gpf_path: str, glg_path: str | None = None, device: int = 0, program: str = "autogrid4",
            bhtree: bool = False, binary_maps: bool = False) -> None:
    """autogrid4 -p gpf -l glg, on one GPU (bhtree: a -DUSE_BHTREE build's
    semantics; binary_maps: also the raw FP32 side-output)."""
    with Context(device) as ctx:
        ctx.run_gpf(gpf_path, glg_path, program, bhtree, binary_maps)


class Scorer:
    """Trilinear ligand scoring against a map set on one GPU (ag_scorer_*).

    ``poses`` is a list of (map index [n], charge [n], position [n, 3] grid
    frame) as poses.make_poses returns.  score() writes the poses' scores
    (and optionally each atom's gradient) into device memory; adjoint()
    writes dL/dmaps = S^T dL/dscores."""

    def __init__(self, ctx: "Context", st: PreparedState, poses, outside: float = 0.0):
        starts = np.zeros(len(poses) + 1, np.int32)
        starts[1:] = np.cumsum([len(p[0]) for p in poses])
        xyz = np.ascontiguousarray(np.concatenate([p[2] for p in poses]), np.float64).reshape(-1, 3)
        mi = np.ascontiguousarray(np.concatenate([p[0] for p in poses]), np.int32)
        q = np.ascontiguousarray(np.concatenate([p[1] for p in poses]), np.float64)
        grid = _abi.Grid((C.c_int32 * 3)(*st.n1), (C.c_int32 * 3)(*st.ne), float(st.spacing))
        h = C.c_void_p()
        check(lib().ag_scorer_create(ctx.handle, C.byref(grid), st.n_maps, st.elec_index, st.dsolv_index,
                                     len(poses), starts.ctypes.data, xyz.ctypes.data, mi.ctypes.data, q.ctypes.data,
                                     float(outside), C.byref(h)))
        self._h = h
        self.ctx = ctx
        self.n_poses = len(poses)
        self.n_atoms = int(starts[-1])
        self.shape = (st.n_maps, st.n1[2], st.n1[1], st.n1[0])

    def set_positions(self, xyz) -> None:
        a = np.ascontiguousarray(xyz, np.float64).reshape(-1)
        assert a.size == 3 * self.n_atoms
        check(lib().ag_scorer_set_positions(self._h, a.ctypes.data))

    def score(self, maps_ptr: int, scores_ptr: int, grad_ptr: int | None = None) -> None:
        check(lib().ag_scorer_score(self._h, C.c_void_p(maps_ptr), C.c_void_p(scores_ptr),
                                    C.c_void_p(grad_ptr) if grad_ptr else None))

    def adjoint(self, dscore_ptr: int, dmaps_ptr: int) -> None:
        check(lib().ag_scorer_adjoint(self._h, C.c_void_p(dscore_ptr), C.c_void_p(dmaps_ptr)))

    def close(self):
        if getattr(self, "_h", None):
            lib().ag_scorer_destroy(self._h)
            self._h = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Analytic:
    """Analytic (differentiable) map model of a state on one GPU
    (ag_analytic_*): forward maps into device memory, adjoint gradients of
    a loss whose cotangent dL/dmaps is given in device memory."""

    def __init__(self, ctx: "Context", st: PreparedState):
        rec, grid, ms, _, setup = st.structs()
        self._keep = list(st._keep)
        w = _abi.Weights(*st.weights)
        h = C.c_void_p()
        check(lib().ag_analytic_create(ctx.handle, C.byref(rec), C.byref(grid), C.byref(ms), C.byref(setup),
                                       C.byref(w), C.byref(h)))
        self._h = h
        self.ctx = ctx
        self.n_atoms = st.n_atoms
        self.n_maps = st.n_maps
        self.shape = (st.n_maps, st.n1[2], st.n1[1], st.n1[0])
        self.points = st.points
        M, T = st.pairs.shape
        self.pair_shape = (M, T)
        self.rij0 = np.ascontiguousarray(st.pairs["Rij"], np.float64)
        wk = np.where(st.pairs["hbonder"] > 0, st.weights[1], st.weights[0])
        self.eps0 = np.ascontiguousarray(st.pairs["eps"] / wk, np.float64)
        self.weights0 = np.array(st.weights, np.float64)
        # Bumped by set() and forward(): the adjoint linearises at the
        # parameters and the tape of the last forward, and a caller that
        # holds an older forward compares this number to know whether that
        # point is still the engine's.
        self.generation = 0
        self.device = "cuda"

    def set(self, coord=None, charge=None, rij=None, eps=None, weights=None) -> None:
        arrs = []

        def ptr(a, n):
            if a is None:
                return None
            a = np.ascontiguousarray(a, np.float64).reshape(-1)
            assert a.size == n
            arrs.append(a)
            return a.ctypes.data
        M, T = self.pair_shape
        w = _abi.Weights(*[float(x) for x in weights]) if weights is not None else None
        check(lib().ag_analytic_set(self._h, ptr(coord, 3 * self.n_atoms), ptr(charge, self.n_atoms),
                                    ptr(rij, M * T), ptr(eps, M * T), C.byref(w) if w is not None else None))
        self.generation += 1

    def forward(self, maps_ptr: int) -> None:
        check(lib().ag_analytic_forward(self._h, C.c_void_p(maps_ptr)))
        self.generation += 1

    def adjoint(self, dmaps_ptr: int) -> dict:
        M, T = self.pair_shape
        g = {"coord": np.zeros((self.n_atoms, 3)), "charge": np.zeros(self.n_atoms),
             "rij": np.zeros((M, T)), "eps": np.zeros((M, T))}
        w = _abi.Weights()
        check(lib().ag_analytic_adjoint(self._h, C.c_void_p(dmaps_ptr), g["coord"].ctypes.data,
                                        g["charge"].ctypes.data, g["rij"].ctypes.data, g["eps"].ctypes.data,
                                        C.byref(w)))
        g["weights"] = np.array([w.vdw, w.hbond, w.estat, w.desolv])
        return g

    def close(self):
        if getattr(self, "_h", None):
            lib().ag_analytic_destroy(self._h)
            self._h = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def build_tables(st: PreparedState):
    """(e_vdw_hb [T,M,1024] float32, sol_fn [8192], r_eps [8192]) of a state's setup (ag_build_tables)."""
    _, _, ms, _, setup = st.structs()
    T, M = setup.n_rec_types, st.n_atom_maps
    ev = np.zeros((T, M, _abi.AG_NEINT), np.float32)
    sol = np.zeros(_abi.AG_NDIEL)
    reps = np.zeros(_abi.AG_NDIEL)
    check(lib().ag_build_tables(C.byref(setup), C.byref(ms), ev.ctypes.data, sol.ctypes.data, reps.ctypes.data))
    return ev, sol, reps


def bond_vectors(st: PreparedState, coord: np.ndarray, type_: np.ndarray, hbond: np.ndarray):
    """(rexp, rvector, rvector2, disorder) of one receptor (ag_bond_vectors)."""
    _, _, _, _, setup = st.structs()
    n = int(coord.shape[0])
    coord = np.ascontiguousarray(coord, np.float64)
    type_ = np.ascontiguousarray(type_, np.int32)
    hbond = np.ascontiguousarray(hbond, np.int32)
    rexp = np.zeros(n, np.int32)
    rv = np.zeros((n, 3))
    rv2 = np.zeros((n, 3))
    dis = np.zeros(n, np.uint8)
    check(lib().ag_bond_vectors(C.byref(setup), n, coord.ctypes.data, type_.ctypes.data, hbond.ctypes.data,
                                rexp.ctypes.data, rv.ctypes.data, rv2.ctypes.data, dis.ctypes.data))
    return rexp, rv, rv2, dis


def bond_vectors_vjp(st: PreparedState, coord, g_rv, g_rv2) -> np.ndarray:
    """Gradient of sum(g_rv . rvector + g_rv2 . rvector2) with respect to the
    coordinates (ag_bond_vectors_vjp)."""
    _, _, _, _, setup = st.structs()
    n = int(coord.shape[0])
    coord = np.ascontiguousarray(coord, np.float64)
    typ = np.ascontiguousarray(st.type, np.int32)
    hb = np.ascontiguousarray(st.hbond, np.int32)
    g_rv = np.ascontiguousarray(g_rv, np.float64)
    g_rv2 = np.ascontiguousarray(g_rv2, np.float64)
    out = np.zeros((n, 3))
    check(lib().ag_bond_vectors_vjp(C.byref(setup), n, coord.ctypes.data, typ.ctypes.data, hb.ctypes.data,
                                    g_rv.ctypes.data, g_rv2.ctypes.data, out.ctypes.data))
    return out


def _with(st: PreparedState, **fields) -> PreparedState:
    """A copy of a state with some fields replaced (dataclasses.replace
    without re-running __init__); its keep list starts empty."""
    new = object.__new__(type(st))
    new.__dict__.update(st.__dict__)
    new.__dict__.update(fields)
    new._keep = []
    return new


def bond_vectors_batch(states: list, n_threads: int = 0) -> list:
    """Copies of the states with bond vectors from ag_bond_vectors_batch
    (host threads); the first state's setup (types, options) applies to all.
    The outputs of all states share one array per field; each state holds
    its slice."""
    n = len(states)
    _, _, _, _, setup = states[0].structs()
    counts = np.array([st.n_atoms for st in states], np.int64)
    offs = np.concatenate([[0], np.cumsum(counts)])
    total = int(offs[-1])
    # every element is written by the library (it clears each output first)
    rexp_all = np.empty(total, np.int32)
    rv_all = np.empty((total, 3))
    rv2_all = np.empty((total, 3))
    dis_all = np.empty(total, np.uint8)
    b_rexp, b_rv, b_rv2, b_dis = (a.ctypes.data for a in (rexp_all, rv_all, rv2_all, dis_all))
    recs = (_abi.ReceptorAddr * n)()
    outs = (_abi.BondOutAddr * n)()
    keep = [list(states[0]._keep)]
    new = []
    for i, st in enumerate(states):
        o0, o1 = int(offs[i]), int(offs[i + 1])
        coord = np.ascontiguousarray(st.coord, np.float64)
        typ = np.ascontiguousarray(st.type, np.int32)
        hb = np.ascontiguousarray(st.hbond, np.int32)
        keep.append((coord, typ, hb))
        r = recs[i]
        r.n_atoms = st.n_atoms
        r.coord, r.type, r.hbond = coord.ctypes.data, typ.ctypes.data, hb.ctypes.data
        ob = outs[i]
        ob.rexp, ob.rvector, ob.rvector2, ob.disorder = b_rexp + 4 * o0, b_rv + 24 * o0, b_rv2 + 24 * o0, b_dis + o0
        new.append(_with(st, rexp=rexp_all[o0:o1], rvector=rv_all[o0:o1], rvector2=rv2_all[o0:o1],
                         disorder=dis_all[o0:o1]))
    check(lib().ag_bond_vectors_batch(C.byref(setup), n, C.cast(recs, C.POINTER(_abi.Receptor)),
                                      C.cast(outs, C.POINTER(_abi.BondOut)), int(n_threads)))
    return new


def xs_radius(type_name: str) -> tuple:
    """(radius, listed) of a type in the xs_vdw table (ag_xs_radius)."""
    r = C.c_double()
    found = lib().ag_xs_radius(type_name.encode(), C.byref(r))
    return float(r.value), bool(found)


def vina_tables(st: PreparedState, lig_types: list, rec_types: list) -> np.ndarray:
    """[T, M, 1024] float32 Vina tables for a state's pairs (ag_vina_tables)."""
    _, _, ms, _, setup = st.structs()
    xs_map = np.array([xs_radius(t)[0] for t in lig_types], np.float64)
    xs_rec = np.array([xs_radius(t)[0] for t in rec_types], np.float64)
    out = np.zeros((len(rec_types), st.n_atom_maps, _abi.AG_NEINT), np.float32)
    check(lib().ag_vina_tables(C.byref(setup), C.byref(ms), xs_map.ctypes.data, xs_rec.ctypes.data,
                               out.ctypes.data))
    return out


def format_map_values(values: np.ndarray) -> bytes:
    """The reference's map text for float32 values (one value per line)."""
    v = np.ascontiguousarray(values, np.float32).reshape(-1)
    cap = 16 * v.size + 16
    buf = C.create_string_buffer(cap)
    n = lib().ag_format_map_values(v.ctypes.data, v.size, buf, cap)
    if n < 0:
        raise AgError(1, "format buffer too small")
    return buf.raw[:n]
