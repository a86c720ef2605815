"""ctypes binding of the gfx950 library (orion_amd/_build/liborion_hip.so).

The library is built in-tree by ``python orion_amd/build.py`` (or
``__graft_entry__.build()``).  Loading fails loudly if it is missing: there
is no CPU fallback for the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import re
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liborion_hip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "orion_hip.h")

FIELD_IDS = {"prim": (0, 6), "metrics": (1, 9), "vol": (2, 1), "q": (3, 5), "dq": (4, 5), "dt": (5, 1),
             "c": (6, 1), "mu": (7, 1)}


class RegionDesc(C.Structure):
    _fields_ = [("block", C.c_int), ("bctype", C.c_int), ("s_st", C.c_int * 3), ("s_ed", C.c_int * 3),
                ("t_st", C.c_int * 3), ("t_ed", C.c_int * 3), ("nbt", C.c_int), ("ibcwin", C.c_int)]


class Inputs(C.Structure):
    _fields_ = [("moo", C.c_double), ("reynolds", C.c_double), ("tref", C.c_double), ("twall", C.c_double),
                ("cfl", C.c_double), ("attack", C.c_double), ("sideslip", C.c_double), ("csrv", C.c_int),
                ("nvis", C.c_int), ("method", C.c_int), ("ntmst", C.c_int), ("gcl", C.c_int)]


class ResidualStats(C.Structure):
    _fields_ = [("rms_residual", C.c_double), ("max_residual", C.c_double), ("max_loc", C.c_int * 5),
                ("phydtime", C.c_double)]

    def as_tuple(self):
        return (self.rms_residual, self.max_residual, tuple(self.max_loc), self.phydtime)


PHASES = ["halo", "bc", "state_dt", "visc_deriv", "visc_face", "visc_diff", "avg1", "inv_face", "inv_closure",
          "lu_pack_state", "inv_diff", "avg2", "lu_pack", "lu_fwd", "lu_bwd", "gs_fwd", "gs_bwd", "lu_chain",
          "update", "residual"]


class OrionHipError(RuntimeError):
    pass


_lib = None
_load_lock = threading.Lock()  # the serving workers may load from several threads


def declared_symbols():
    """Function names declared in include/orion_hip.h."""
    with open(HEADER) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(orion_hip_[a-z_]+)\s*\(", text)))


def load():
    global _lib
    if _lib is not None:
        return _lib
    with _load_lock:
        if _lib is None:
            _lib = _open()
    return _lib


def _open():
    if not os.path.exists(LIB_PATH):
        raise OrionHipError("HIP library not built: %s is missing (run python orion_amd/build.py)" % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    P = C.c_void_p
    sig = {
        "orion_hip_backend_init": ([C.c_int], C.c_int),
        "orion_hip_last_error": ([], C.c_char_p),
        "orion_hip_ctx_create": ([C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double), C.c_int,
                                  C.POINTER(RegionDesc), C.POINTER(Inputs), C.c_int, C.POINTER(P)], C.c_int),
        "orion_hip_batch_create": ([C.c_int, C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double),
                                   C.c_int, C.POINTER(RegionDesc), C.POINTER(Inputs), C.c_int, C.POINTER(P)], C.c_int),
        "orion_hip_batch_create_grids": ([C.c_int, C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_int),
                                          C.POINTER(C.POINTER(C.c_double)), C.c_int, C.POINTER(RegionDesc),
                                          C.POINTER(Inputs), C.c_int, C.POINTER(P)], C.c_int),
        "orion_hip_requests": ([P], C.c_int),
        "orion_hip_ctx_destroy": ([P], None),
        "orion_hip_field": ([P, C.c_int, C.c_int, C.POINTER(C.c_void_p), C.POINTER(C.c_longlong)], C.c_int),
        "orion_hip_upload": ([P, C.c_int, C.c_int, C.POINTER(C.c_double), C.c_longlong], C.c_int),
        "orion_hip_download": ([P, C.c_int, C.c_int, C.POINTER(C.c_double), C.c_longlong], C.c_int),
        "orion_hip_copy_block_state": ([P, C.c_int, P, C.c_int], C.c_int),
        "orion_hip_set_per_block": ([P, C.c_int], C.c_int),
        "orion_hip_set_visc_tiles": ([P, C.c_int, C.c_int], C.c_int),
        "orion_hip_get_visc_tiles": ([P, C.POINTER(C.c_int), C.POINTER(C.c_int)], C.c_int),
        "orion_hip_set_visc_form": ([P, C.c_int], C.c_int),
        "orion_hip_get_visc_form": ([P], C.c_int),
        "orion_hip_tune": ([P, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                            C.POINTER(C.c_double), C.POINTER(C.c_int)], C.c_int),
        "orion_hip_residual_noalloc": ([C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                        C.POINTER(C.c_void_p), C.c_void_p, C.POINTER(ResidualStats)], C.c_int),
        "orion_hip_update_temperature": ([P], C.c_int),
        "orion_hip_exchange_bc": ([P], C.c_int),
        "orion_hip_apply_physical_bc": ([P], C.c_int),
        "orion_hip_update_flow_states": ([P], C.c_int),
        "orion_hip_calculate_time_step": ([P, C.POINTER(C.c_double)], C.c_int),
        "orion_hip_compute_viscous_rhs": ([P], C.c_int),
        "orion_hip_average_interface_residuals": ([P], C.c_int),
        "orion_hip_compute_inviscid_rhs": ([P], C.c_int),
        "orion_hip_update_solution": ([P], C.c_int),
        "orion_hip_residual": ([P, C.POINTER(ResidualStats)], C.c_int),
        "orion_hip_run": ([P, C.c_int, C.POINTER(ResidualStats)], C.c_int),
        "orion_hip_profile_run": ([P, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_longlong)], C.c_int),
        "orion_hip_synchronize": ([P], C.c_int),
        "orion_hip_set_stream": ([P, C.c_void_p], C.c_int),
        "orion_hip_set_overlap": ([P, C.c_int], C.c_int),
        "orion_hip_set_schedule": ([P, C.c_int], C.c_int),
        "orion_hip_get_schedule": ([P], C.c_int),
        "orion_hip_tune_schedule": ([P, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double)], C.c_int),
        "orion_hip_device_bytes": ([P], C.c_size_t),
        "orion_hip_interior_nodes": ([P], C.c_longlong),
        "orion_hip_output_fields": ([P, C.c_int, C.POINTER(C.c_float), C.c_longlong], C.c_int),
        "orion_hip_write_solution": ([P, C.c_char_p, C.c_int, C.POINTER(C.c_double)], C.c_int),
        "orion_hip_write_vts": ([C.c_char_p, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double),
                                 C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_float)], C.c_int),
        "orion_hip_write_vtm": ([C.c_char_p, C.c_char_p, C.c_int, C.c_int], C.c_int),
        "orion_hip_save_state": ([P, C.c_char_p, C.c_longlong], C.c_int),
        "orion_hip_load_state": ([P, C.c_char_p, C.POINTER(C.c_longlong)], C.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    return L


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = load().orion_hip_last_error().decode(errors="replace")
        raise OrionHipError("%s: %s" % (what, msg) if what else msg)
