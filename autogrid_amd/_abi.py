"""ctypes binding of the C-ABI declared in include/autogrid_amd.h.

The engine is the in-tree shared library ``autogrid_amd/libautogrid_amd.so``
(C++ host front end + gfx950 HIP kernels).  Loading it fails loudly when it
has not been built; there is no Python or CPU fallback for the map
computation.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# AUTOGRID_AMD_LIB / AUTOGRID_AMD_CLI select another build of the same
# sources (the host-sanitizer build of `make -C autogrid_amd sanitize`).
LIB_PATH = os.environ.get("AUTOGRID_AMD_LIB") or os.path.join(_HERE, "libautogrid_amd.so")
CLI_PATH = os.environ.get("AUTOGRID_AMD_CLI") or os.path.join(_HERE, "bin", "autogrid4")

AG_OK = 0
STATUS_NAMES = {0: "AG_OK", 1: "AG_ERR_BAD_INPUT", 2: "AG_ERR_NO_DONOR_H", 3: "AG_ERR_OOM",
                4: "AG_ERR_DEVICE", 5: "AG_ERR_IO", 6: "AG_ERR_UNSUPPORTED"}
AG_JOB_NO_FILES = 1
AG_JOB_BHTREE = 2
AG_JOB_BINARY_MAPS = 4
AG_NEINT = 1024
AG_POTENTIAL_AD4 = 0
AG_POTENTIAL_VINA = 1
AG_NDIEL = 8192
AG_ERR_UNSUPPORTED = 6
AG_NOALLOC_SCRATCH_BYTES = 32 << 20    # include/autogrid_amd.h


class AgError(RuntimeError):
    """A non-OK ag_status, with the library's message."""

    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


class Receptor(C.Structure):
    _fields_ = [("n_atoms", C.c_int32), ("coord", C.POINTER(C.c_double)),
                ("charge", C.POINTER(C.c_double)), ("vol", C.POINTER(C.c_double)),
                ("solpar", C.POINTER(C.c_double)), ("type", C.POINTER(C.c_int32)),
                ("hbond", C.POINTER(C.c_int32)), ("rexp", C.POINTER(C.c_int32)),
                ("rvector", C.POINTER(C.c_double)), ("rvector2", C.POINTER(C.c_double)),
                ("disorder", C.POINTER(C.c_uint8))]


class ReceptorAddr(C.Structure):
    """Receptor with its pointers as plain addresses (same layout), filled
    from ndarray.ctypes.data for many conformations at once."""
    _fields_ = [("n_atoms", C.c_int32)] + [(name, C.c_void_p) for name, _ in Receptor._fields_[1:]]


assert C.sizeof(ReceptorAddr) == C.sizeof(Receptor)


class Grid(C.Structure):
    _fields_ = [("n1", C.c_int32 * 3), ("ne", C.c_int32 * 3), ("spacing", C.c_double)]


class MapDesc(C.Structure):
    _fields_ = [("is_covalent", C.c_int32), ("is_hbonder", C.c_int32), ("hbond", C.c_int32),
                ("reserved", C.c_int32), ("solpar_probe", C.c_double), ("vol_probe", C.c_double)]


class MapSet(C.Structure):
    _fields_ = [("n_atom_maps", C.c_int32), ("maps", C.POINTER(MapDesc)),
                ("elec_index", C.c_int32), ("dsolv_index", C.c_int32), ("n_maps", C.c_int32),
                ("separate_desolvation", C.c_int32), ("disorder_h", C.c_int32),
                ("hydrogen_type", C.c_int32), ("dddiel", C.c_int32), ("floating_grid", C.c_int32),
                ("constriction_grid", C.c_int32), ("bhtree", C.c_int32),
                ("inv_diel_cal", C.c_double), ("coeff_estat", C.c_double), ("solpar_q", C.c_double),
                ("cov_pos", C.c_double * 3), ("cov_half_width", C.c_double),
                ("cov_barrier", C.c_double), ("constriction_cutoff", C.c_double)]


class Tables(C.Structure):
    _fields_ = [("n_rec_types", C.c_int32), ("reserved", C.c_int32),
                ("e_vdw_hb", C.POINTER(C.c_float)), ("sol_fn", C.POINTER(C.c_double)),
                ("r_eps", C.POINTER(C.c_double))]


class Pair(C.Structure):
    _fields_ = [("cA", C.c_double), ("cB", C.c_double), ("Rij", C.c_double), ("eps", C.c_double),
                ("xA", C.c_int32), ("xB", C.c_int32), ("hbonder", C.c_int32), ("reserved", C.c_int32)]


class Tune(C.Structure):
    _fields_ = [("cell_size", C.c_double), ("cube", C.c_int32), ("reserved", C.c_int32), ("tune_ms", C.c_double)]


class Weights(C.Structure):
    _fields_ = [("vdw", C.c_double), ("hbond", C.c_double), ("estat", C.c_double), ("desolv", C.c_double)]


class BondOut(C.Structure):
    _fields_ = [("rexp", C.POINTER(C.c_int32)), ("rvector", C.POINTER(C.c_double)),
                ("rvector2", C.POINTER(C.c_double)), ("disorder", C.POINTER(C.c_uint8))]


class BondOutAddr(C.Structure):
    """BondOut with its pointers as plain addresses (same layout)."""
    _fields_ = [(name, C.c_void_p) for name, _ in BondOut._fields_]


assert C.sizeof(BondOutAddr) == C.sizeof(BondOut)


class Setup(C.Structure):
    _fields_ = [("n_rec_types", C.c_int32), ("n_atom_maps", C.c_int32),
                ("pairs", C.POINTER(Pair)), ("r_smooth", C.c_double), ("coeff_desolv", C.c_double),
                ("dddiel", C.c_int32), ("disorder_h", C.c_int32),
                ("t_hydrogen", C.c_int32), ("t_nonhb_hydrogen", C.c_int32), ("t_carbon", C.c_int32),
                ("t_arom_carbon", C.c_int32), ("t_oxygen", C.c_int32), ("t_sulphur", C.c_int32),
                ("t_nitrogen", C.c_int32), ("t_nonhb_nitrogen", C.c_int32)]


_lib = None


def lib() -> C.CDLL:
    """The loaded engine library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                               "(the engine has no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.ag_last_error.restype = C.c_char_p
        L.ag_abi_version.restype = C.c_int
        L.ag_context_create.argtypes = [C.POINTER(vp), C.c_int]
        L.ag_context_destroy.argtypes = [vp]
        L.ag_job_open.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_uint32, C.POINTER(vp)]
        L.ag_job_prepare.argtypes = [vp]
        L.ag_job_views.argtypes = [vp, C.POINTER(Receptor), C.POINTER(Grid), C.POINTER(MapSet),
                                   C.POINTER(Tables), C.POINTER(Setup)]
        L.ag_job_points.argtypes = [vp]
        L.ag_job_points.restype = C.c_int64
        L.ag_job_close.argtypes = [vp]
        L.ag_job_close.restype = None
        L.ag_run_gpf.argtypes = [vp, C.c_char_p, C.c_char_p, C.c_char_p]
        L.ag_run_gpf_flags.argtypes = [vp, C.c_char_p, C.c_char_p, C.c_char_p, C.c_uint32]
        L.ag_run_gpf_flags.restype = C.c_int
        L.ag_maps_plain.argtypes = [vp, C.POINTER(Receptor), C.POINTER(Grid), C.POINTER(MapSet),
                                    C.POINTER(Tables), vp, vp, vp, C.c_int]
        L.ag_format_map_values.argtypes = [vp, C.c_int64, C.c_char_p, C.c_int64]
        L.ag_format_map_values.restype = C.c_int64
        L.ag_job_write_maps.argtypes = [vp, vp, vp, vp, C.c_int]
        L.ag_plan_create.argtypes = [vp, C.POINTER(Receptor), C.POINTER(Grid), C.POINTER(MapSet),
                                     C.POINTER(Tables), C.POINTER(vp)]
        L.ag_plan_execute.argtypes = [vp, vp, vp, vp]
        L.ag_plan_execute_host.argtypes = [vp, vp, vp, vp, vp, vp, vp]
        L.ag_plan_destroy.argtypes = [vp]
        L.ag_plan_audit.argtypes = [vp, C.POINTER(C.c_int64 * 4)]
        L.ag_context_timing.argtypes = [vp, C.POINTER(C.c_float)]
        L.ag_xs_radius.argtypes = [C.c_char_p, C.POINTER(C.c_double)]
        L.ag_vina_tables.argtypes = [C.POINTER(Setup), C.POINTER(MapSet), vp, vp, vp]
        L.ag_job_prepare_vina.argtypes = [vp]
        L.ag_maps_vina.argtypes = L.ag_maps_plain.argtypes
        L.ag_maps_plain_noalloc.argtypes = [vp, C.POINTER(Receptor), C.POINTER(Grid), C.POINTER(MapSet),
                                            C.POINTER(Tables), vp, vp, vp]
        L.ag_maps_plain_noalloc.restype = C.c_int
        L.ag_plan_create_potential.argtypes = [vp, C.POINTER(Receptor), C.POINTER(Grid), C.POINTER(MapSet),
                                               C.POINTER(Tables), C.c_int, C.POINTER(vp)]
        L.ag_build_tables.argtypes = [C.POINTER(Setup), C.POINTER(MapSet), vp, vp, vp]
        L.ag_bond_vectors.argtypes = [C.POINTER(Setup), C.c_int, vp, vp, vp, vp, vp, vp, vp]
        L.ag_batch_create.argtypes = [vp, C.c_int, C.POINTER(Receptor), C.POINTER(Grid), C.POINTER(MapSet),
                                      C.POINTER(Tables), C.c_int, C.POINTER(Tune), C.POINTER(vp)]
        L.ag_plan_layout.argtypes = [vp, vp, vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.ag_plan_conformations.argtypes = [vp]
        L.ag_plan_conformations.restype = C.c_int
        L.ag_plan_tune.argtypes = [vp, C.POINTER(Tune)]
        L.ag_maps_batch.argtypes = [vp, C.c_int, C.POINTER(Receptor), C.POINTER(Grid), C.POINTER(MapSet),
                                    C.POINTER(Tables), vp, vp, vp, C.c_int, C.POINTER(Tune)]
        L.ag_bond_vectors_batch.argtypes = [C.POINTER(Setup), C.c_int, C.POINTER(Receptor), C.POINTER(BondOut),
                                            C.c_int]
        L.ag_analytic_create.argtypes = [vp, C.POINTER(Receptor), C.POINTER(Grid), C.POINTER(MapSet),
                                         C.POINTER(Setup), C.POINTER(Weights), C.POINTER(vp)]
        L.ag_analytic_set.argtypes = [vp, vp, vp, vp, vp, C.POINTER(Weights)]
        L.ag_analytic_forward.argtypes = [vp, vp]
        L.ag_analytic_adjoint.argtypes = [vp, vp, vp, vp, vp, vp, C.POINTER(Weights)]
        L.ag_analytic_destroy.argtypes = [vp]
        L.ag_job_weights.argtypes = [vp, C.POINTER(Weights)]
        L.ag_scorer_create.argtypes = [vp, C.POINTER(Grid), C.c_int, C.c_int, C.c_int, C.c_int, vp, vp, vp, vp,
                                       C.c_double, C.POINTER(vp)]
        L.ag_scorer_set_positions.argtypes = [vp, vp]
        L.ag_scorer_score.argtypes = [vp, vp, vp, vp]
        L.ag_scorer_adjoint.argtypes = [vp, vp, vp]
        L.ag_scorer_destroy.argtypes = [vp]
        L.ag_bond_vectors_vjp.argtypes = [C.POINTER(Setup), C.c_int, vp, vp, vp, vp, vp, vp]
        L.ag_bond_vectors_vjp.restype = C.c_int
        for name in ("ag_build_tables", "ag_bond_vectors", "ag_batch_create", "ag_plan_layout", "ag_plan_tune",
                     "ag_maps_batch", "ag_bond_vectors_batch", "ag_analytic_create", "ag_analytic_set",
                     "ag_analytic_forward", "ag_analytic_adjoint", "ag_analytic_destroy", "ag_job_weights",
                     "ag_scorer_create", "ag_scorer_set_positions", "ag_scorer_score", "ag_scorer_adjoint",
                     "ag_scorer_destroy"):
            getattr(L, name).restype = C.c_int
        for name in ("ag_plan_create", "ag_plan_execute", "ag_plan_execute_host", "ag_plan_destroy", "ag_plan_audit", "ag_context_timing",
                     "ag_context_create", "ag_context_destroy", "ag_job_open", "ag_job_prepare",
                     "ag_job_views", "ag_run_gpf", "ag_maps_plain", "ag_job_write_maps", "ag_xs_radius",
                     "ag_vina_tables", "ag_job_prepare_vina", "ag_maps_vina", "ag_plan_create_potential"):
            getattr(L, name).restype = C.c_int
        _lib = L
    return _lib


def check(status: int) -> None:
    if status != AG_OK:
        raise AgError(status, lib().ag_last_error().decode(errors="replace"))


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))
