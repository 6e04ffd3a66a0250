"""ctypes binding of libsgeo.so (include/sgeo.h).

The HIP library is the only compute path: if it is missing or fails to load
this module raises, and nothing falls back to a CPU implementation.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsgeo.so")
_lib = None

d_p = C.POINTER(C.c_double)
f_p = C.POINTER(C.c_float)
i_p = C.POINTER(C.c_int32)
u8_p = C.POINTER(C.c_uint8)
vp = C.c_void_p

# name -> (restype, argtypes); mirrors include/sgeo.h
SIGNATURES = {
    "sgeo_init": (C.c_int, [C.c_int, C.POINTER(vp)]),
    "sgeo_finalize": (C.c_int, [vp]),
    "sgeo_last_error": (C.c_char_p, []),
    "sgeo_device_info": (C.c_int, [vp, C.c_char_p, C.c_size_t, C.POINTER(C.c_int),
                                   C.POINTER(C.c_size_t)]),
    "sgeo_synchronize": (C.c_int, [vp]),
    "sgeo_set_reduction": (C.c_int, [vp, C.c_int]),
    "sgeo_profile": (C.c_int, [vp, C.c_int, C.POINTER(C.c_int64),
                               C.POINTER(C.c_double)]),
    "sgeo_pcg_solver": (C.c_int, [vp, C.c_int32, C.c_int32, d_p, d_p, d_p, d_p, i_p,
                                  C.c_double, C.c_int32, i_p, i_p, C.c_char_p]),
    "sgeo_pcg_solver_invalidate": (C.c_int, [vp]),
    "sgeo_pcg_solver_noalloc": (C.c_int, [vp, C.c_int32, C.c_int32, vp, vp, vp, vp, vp,
                                          C.c_double, C.c_int32, i_p, i_p, C.c_char_p]),
    "sgeo_pcg_solver_noalloc_bytes": (C.c_size_t, [C.c_int32, C.c_int32]),
    "sgeo_mesh_create": (C.c_int, [vp, C.c_int32, C.c_int32, d_p, i_p, i_p,
                                   C.c_int32, C.POINTER(vp)]),
    "sgeo_mesh_destroy": (C.c_int, [vp]),
    "sgeo_mesh_build_stiffness": (C.c_int, [vp, d_p, d_p, d_p, C.c_int, d_p, d_p]),
    "sgeo_mesh_upload_stiffness": (C.c_int, [vp, d_p, d_p, C.c_int]),
    "sgeo_mesh_download_stiffness": (C.c_int, [vp, d_p]),
    "sgeo_mesh_set_operator": (C.c_int, [vp, C.c_int, C.c_int]),
    "sgeo_calibrate_copy": (C.c_int, [vp, C.c_size_t, C.c_int, d_p]),
    "sgeo_mesh_matvec": (C.c_int, [vp, d_p, d_p]),
    "sgeo_constitutive": (C.c_int, [vp, C.c_int32, d_p, d_p]),
    "sgeo_math": (C.c_int, [vp, C.c_int32, C.c_int32, d_p, d_p]),
    "sgeo_mesh_bytes": (C.c_int, [vp, C.POINTER(C.c_size_t)]),
    "sgeo_slope_create": (C.c_int, [vp, d_p, C.POINTER(vp)]),
    "sgeo_slope_destroy": (C.c_int, [vp]),
    "sgeo_slope_begin_srf": (C.c_int, [vp, d_p, d_p, d_p, d_p, d_p, C.c_double]),
    "sgeo_slope_plastic_iter": (C.c_int, [vp, C.c_double, C.c_int32, C.c_double,
                                          C.c_int, i_p, i_p, d_p, i_p, d_p]),
    "sgeo_slope_begin_increment": (C.c_int, [vp, d_p, d_p, d_p, d_p, d_p, C.c_double,
                                             d_p]),
    "sgeo_slope_download_vmeps": (C.c_int, [vp, d_p]),
    "sgeo_mesh_elastic_stress": (C.c_int, [vp, d_p, d_p, d_p, d_p]),
    "sgeo_mesh_solve": (C.c_int, [vp, d_p, d_p, C.c_double, C.c_int32, i_p, i_p]),
    "sgeo_mesh_excavation_load": (C.c_int, [vp, d_p, d_p, d_p]),
    "sgeo_slope_run_srf": (C.c_int, [vp, i_p, C.c_int32, C.c_double, C.c_int32,
                                     C.c_double, C.c_int32, i_p, i_p, d_p, d_p,
                                     d_p, i_p]),
    "sgeo_slope_download": (C.c_int, [vp, d_p, d_p, d_p, d_p, u8_p]),
    "sgeo_batch_create": (C.c_int, [vp, C.c_int32, d_p, C.POINTER(vp)]),
    "sgeo_batch_destroy": (C.c_int, [vp]),
    "sgeo_batch_bytes": (C.c_int, [vp, C.POINTER(C.c_size_t)]),
    "sgeo_batch_begin": (C.c_int, [vp, C.c_int32, d_p, d_p, d_p, d_p, d_p, C.c_double,
                                   d_p, C.c_int32]),
    "sgeo_batch_release": (C.c_int, [vp, C.c_int32]),
    "sgeo_batch_plastic_iter": (C.c_int, [vp, i_p, C.c_double, C.c_int32, C.c_double,
                                          C.c_double, i_p, i_p, d_p, i_p, d_p]),
    "sgeo_batch_advance": (C.c_int, [vp, C.c_double, C.c_int32, C.c_double, C.c_double,
                                     C.c_int32, C.c_int64, i_p, i_p, i_p, d_p, i_p,
                                     i_p]),
    "sgeo_batch_download": (C.c_int, [vp, C.c_int32, d_p, u8_p]),
    "sgeo_hex2spec": (C.c_int, [vp, C.c_int32, C.c_int32, vp, vp, C.c_int32, vp,
                                C.POINTER(vp), C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "sgeo_spec_fetch": (C.c_int, [vp, vp, vp, vp, vp, vp]),
    "sgeo_spec_destroy": (C.c_int, [vp]),
    "sgeo_batch_set_graphs": (C.c_int, [vp, C.c_int]),
    "sgeo_mesh_set_nodes": (C.c_int, [vp, C.c_int32, C.POINTER(C.c_int32)]),
    "sgeo_slope_nodal_fields": (C.c_int, [vp, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "sgeo_batch_nodal_fields": (C.c_int, [vp, C.c_int32, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "sgeo_slope_accumulate_nodal": (C.c_int, [vp, C.c_int32]),
    "sgeo_batch_begin_increment": (C.c_int, [vp, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double),
                                             C.c_double, C.POINTER(C.c_double)]),
    "sgeo_batch_download_stress": (C.c_int, [vp, C.c_int32, C.POINTER(C.c_double)]),
    "sgeo_batch_accumulate_nodal": (C.c_int, [vp, C.c_int32, C.c_int32]),
    "sgeo_mesh_nodal_fields": (C.c_int, [vp, C.POINTER(C.c_double), C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "sgeo_batch_nl_iters": (C.c_int, [vp, C.POINTER(C.c_int32)]),
    "sgeo_batch_set_geometry": (C.c_int, [vp, C.c_int32, C.c_int32]),
    "sgeo_batch_set_residual_period": (C.c_int, [vp, C.c_int32]),
    "sgeo_batch_set_deferred_x": (C.c_int, [vp, C.c_int32]),
    "sgeo_batch_autotune": (C.c_int, [vp, C.c_int32, C.POINTER(C.c_int32),
                                      C.POINTER(C.c_int32), C.POINTER(C.c_double),
                                      C.POINTER(C.c_double)]),
    "sgeo_batch_stats": (C.c_int, [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                   C.POINTER(C.c_int64)]),
}


class SgeoError(RuntimeError):
    pass


def lib():
    """Load libsgeo.so (raises if it is absent: there is no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SgeoError(f"{LIB_PATH} not built: run "
                            "`python -m specfem3d_geotech_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int, what: str = ""):
    if rc != 0:
        msg = lib().sgeo_last_error().decode(errors="replace")
        raise SgeoError(f"{what} failed ({rc}): {msg}")


def ptr(a: np.ndarray | None, ct=C.c_double):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ct))


def f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)
