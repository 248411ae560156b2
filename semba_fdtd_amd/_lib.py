"""ctypes binding of libsembafdtd.so (the C ABI declared in include/sembafdtd.h).

The library is loaded from this package directory only; there is no CPU
fallback.  Importing works without a GPU (for symbol checks); creating a grid
without a gfx950 device raises ``FdtdError``.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsembafdtd.so")

f32p = C.POINTER(C.c_float)
f64p = C.POINTER(C.c_double)
i32p = C.POINTER(C.c_int32)
i16p = C.POINTER(C.c_int16)
vp = C.c_void_p


class FdtdGridDesc(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("mode", C.c_int32), ("storage", C.c_int32), ("device", C.c_int32),
                ("stream", vp)]


MODE_FAST, MODE_EXACT = 0, 1
STORE_FP32, STORE_FP16 = 0, 1

# name -> (restype, argtypes)
SIGNATURES = {
    "fdtd_abi_version": (C.c_int, []),
    "fdtd_create": (C.c_int, [C.POINTER(FdtdGridDesc), C.POINTER(vp)]),
    "fdtd_destroy": (C.c_int, [vp]),
    "fdtd_last_error": (C.c_char_p, [vp]),
    "fdtd_set_media": (C.c_int, [vp, C.c_int32, f32p, f32p, f32p, f32p, i32p]),
    "fdtd_set_media_index": (C.c_int, [vp, C.c_int32, i16p]),
    "fdtd_set_spacing": (C.c_int, [vp] + [f32p] * 6),
    "fdtd_set_cpml": (C.c_int, [vp, i32p] + [f32p] * 12),
    "fdtd_set_edispersive": (C.c_int, [vp, C.c_int32, i32p, i32p, f64p, f64p, f64p]),
    "fdtd_set_mdispersive": (C.c_int, [vp, C.c_int32, i32p, i32p, f64p, f64p, f64p]),
    "fdtd_set_time": (C.c_int, [vp, C.c_double, C.c_int32, f64p]),
    "fdtd_add_nodal": (C.c_int, [vp, C.c_int32, C.c_int32, i32p, C.c_float, C.c_float,
                                 C.c_int32, f32p]),
    "fdtd_advance_e": (C.c_int, [vp, C.c_int32]),
    "fdtd_advance_h": (C.c_int, [vp, C.c_int32]),
    "fdtd_advance_e_yee": (C.c_int, [vp]),
    "fdtd_advance_e_cpml": (C.c_int, [vp]),
    "fdtd_advance_e_dispersive": (C.c_int, [vp]),
    "fdtd_advance_planewave_e": (C.c_int, [vp, C.c_int32]),
    "fdtd_advance_planewave_h": (C.c_int, [vp, C.c_int32]),
    "fdtd_set_lines": (C.c_int, [vp, f32p, f32p, f32p]),
    "fdtd_set_tiling": (C.c_int, [vp, C.c_int32, C.c_int32, C.c_int32]),
    "fdtd_state_size": (C.c_int64, [vp]),
    "fdtd_set_field_scale": (C.c_int, [vp, C.c_int32, C.c_float]),
    "fdtd_save_state": (C.c_int, [vp, C.c_int32, vp, C.c_int64]),
    "fdtd_load_state": (C.c_int, [vp, vp, C.c_int64, i32p]),
    "fdtd_batch_create": (C.c_int, [C.c_int32, vp, C.POINTER(vp)]),
    "fdtd_batch_destroy": (C.c_int, [vp]),
    "fdtd_batch_last_error": (C.c_char_p, [vp]),
    "fdtd_batch_admit": (C.c_int, [vp, vp, C.c_int32, i32p]),
    "fdtd_batch_retire": (C.c_int, [vp, C.c_int32]),
    "fdtd_batch_step_of": (C.c_int64, [vp, C.c_int32]),
    "fdtd_batch_set_tiling": (C.c_int, [vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32]),
    "fdtd_batch_set_step_mode": (C.c_int, [vp, C.c_int32, C.c_int32]),
    "fdtd_batch_step": (C.c_int, [vp, C.c_int32]),
    "fdtd_batch_synchronize": (C.c_int, [vp]),
    "fdtd_add_dft": (C.c_int, [vp, C.c_int32, i32p, i32p, C.c_int32, f32p, i32p]),
    "fdtd_dft_size": (C.c_int64, [vp, C.c_int32]),
    "fdtd_read_dft": (C.c_int, [vp, C.c_int32, f64p, C.c_int64]),
    "fdtd_update_observation": (C.c_int, [vp, C.c_int32]),
    "fdtd_get_tiling": (C.c_int, [vp, i32p, i32p, i32p]),
    "fdtd_time_halves": (C.c_int, [vp, C.c_int32, f32p]),
    "fdtd_set_step_mode": (C.c_int, [vp, C.c_int32, C.c_int32]),
    "fdtd_set_graph_steps": (C.c_int, [vp, C.c_int32]),
    "fdtd_get_step_mode": (C.c_int, [vp, i32p, i32p]),
    "fdtd_set_step_class_rows": (C.c_int, [vp, C.c_int32, C.c_int32]),
    "fdtd_get_step_class_rows": (C.c_int, [vp, i32p]),
    "fdtd_set_step_vector": (C.c_int, [vp, C.c_int32]),
    "fdtd_time_step": (C.c_int, [vp, C.c_int32, f32p]),
    "fdtd_step_classes": (C.c_int, [vp, C.c_int32, i32p, f64p, f64p, f32p]),
    "fdtd_add_planewave": (C.c_int, [vp, i32p, C.c_int32, f32p, f32p, f32p, C.c_float,
                                     C.c_int32, f32p, C.c_float, C.c_float]),
    "fdtd_build_lines": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, f32p, f32p]),
    "fdtd_advance_nodal_e": (C.c_int, [vp, C.c_int32]),
    "fdtd_advance_h_yee": (C.c_int, [vp]),
    "fdtd_advance_h_cpml": (C.c_int, [vp]),
    "fdtd_advance_h_dispersive": (C.c_int, [vp]),
    "fdtd_advance_mur": (C.c_int, [vp]),
    "fdtd_set_mur": (C.c_int, [vp, i32p, f32p]),
    "fdtd_advance_nodal_h": (C.c_int, [vp, C.c_int32]),
    "fdtd_run": (C.c_int, [vp, C.c_int32, C.c_int32]),
    "fdtd_set_probes": (C.c_int, [vp, C.c_int32, i32p, i32p, i32p, i32p, C.c_int32]),
    "fdtd_read_probes": (C.c_int, [vp, f32p, C.c_int32, i32p]),
    "fdtd_sample_probes": (C.c_int, [vp]),
    "fdtd_get_field": (C.c_int, [vp, C.c_int32, f32p]),
    "fdtd_set_field": (C.c_int, [vp, C.c_int32, f32p]),
    "fdtd_get_field_noalloc": (C.c_int, [vp, C.c_int32, f32p]),
    "fdtd_synchronize": (C.c_int, [vp]),
    "fdtd_device_bytes": (C.c_int64, [vp]),
    "fdtd_device_memory": (C.c_int, [C.c_int32, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "fdtd_dispersive_edges": (C.c_int64, [vp, C.c_int32]),
    "fdtd_courant_dt": (C.c_double, [C.c_float] * 6 + [C.c_int32, C.c_double]),
    "fdtd_build_g_tables": (C.c_int, [C.c_int32] + [f32p] * 4 + [i32p, C.c_double,
                                       C.c_float, C.c_float] + [f32p] * 4),
    "fdtd_build_spacing": (C.c_int, [C.c_int32] + [f32p] * 5),
    "fdtd_build_cpml_axis": (C.c_int, [C.c_int32] * 3 + [C.c_float] * 7 +
                             [C.c_double, C.c_float, C.c_float] + [f32p] * 8),
    "fdtd_build_ade": (C.c_int, [C.c_int32, f64p, f64p, C.c_float, C.c_float, C.c_double,
                                 f64p, f64p, f64p, f32p, f32p]),
    "fdtd_build_mur": (C.c_int, [C.c_int32, f32p, f32p, C.c_double, C.c_float, C.c_float, f32p,
                                 f32p]),
    "fdtd_evolucion": (C.c_float, [C.c_float, C.c_float, C.c_int32, f32p]),
    "fdtd_media_fill": (C.c_int, [vp, C.c_int16]),
    "fdtd_media_pec_walls": (C.c_int, [vp, C.c_int32]),
    "fdtd_media_paint_sphere": (C.c_int, [vp] + [C.c_float] * 4 + [C.c_int16, C.c_int32, i32p]),
    "fdtd_media_paint_box": (C.c_int, [vp] + [C.c_int32] * 6 + [C.c_int16, C.c_int32, i32p]),
}


class FdtdError(RuntimeError):
    pass


_LIB = None


def load(path: str = LIB_PATH):
    """Load the library and bind every entry point; raises if it is missing."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise FdtdError(f"{path} is missing: build it with "
                        f"`python -c 'import __graft_entry__ as g; g.build()'`")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


def ptr(a, t):
    return a.ctypes.data_as(t)
