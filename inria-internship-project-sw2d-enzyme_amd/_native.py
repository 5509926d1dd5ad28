"""ctypes bindings of the native libraries built in ``lib/``.

``libsw2dg.so``      -- HIP kernels + C-ABI of include/sw2dg.h (and the host tooling)
``libsw2d_host.so``  -- host tooling only (include/sw2d_host.h); loads without a GPU

Nothing here falls back to Python: a missing library raises ImportError-like
errors with the build command to run.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(HERE, "lib")

# status codes (include/sw2dg.h)
OK, ERR_ARG, ERR_NAN, ERR_NEG_DEPTH, ERR_FLOW_CONFIG, ERR_BC, ERR_HIP, ERR_NO_DEVICE = range(8)
GODUNOV, MUSCL_EVR = 0, 1
TOPO_DISCONTINUOUS, TOPO_BLENDING = 0, 1
FRIC_NONE, FRIC_MANNING, FRIC_STRICKLER, FRIC_CHEZY = 0, 1, 2, 3
BC_FR, BC_H, BC_Z, BC_Q = 0, 1, 2, 3
BC_TYPES = {"fr": BC_FR, "h": BC_H, "z": BC_Z, "q": BC_Q}


class SW2DError(RuntimeError):
    """Error status returned by the native library."""

    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


_d = C.POINTER(C.c_double)
_i64 = C.POINTER(C.c_int64)
_i32 = C.POINTER(C.c_int32)


class Mesh(C.Structure):
    """sw2dg_mesh (include/sw2dg.h)."""

    _fields_ = [
        ("n_cell", C.c_int64), ("n_bnd", C.c_int64), ("n_int", C.c_int64),
        ("cell_area", _d), ("cell_x", _d), ("cell_y", _d), ("cell_zb", _d),
        ("bnd_cell", _i64), ("bnd_len", _d), ("bnd_nx", _d), ("bnd_ny", _d),
        ("bnd_series", _i32),
        ("int_L", _i64), ("int_R", _i64),
        ("int_len", _d), ("int_nx", _d), ("int_ny", _d), ("int_xm", _d), ("int_ym", _d),
    ]


class Params(C.Structure):
    """sw2dg_params (include/sw2dg.h)."""

    _fields_ = [
        ("scheme", C.c_int32), ("topo", C.c_int32), ("friction", C.c_int32),
        ("head_loss", C.c_int32), ("neg_depth_detection", C.c_int32),
        ("h_min", C.c_double), ("phi_min", C.c_double), ("cfl_max", C.c_double),
        ("dt_max", C.c_double), ("dt_map", C.c_double), ("dt_time_series", C.c_double),
        ("t0", C.c_double), ("t_max", C.c_double), ("v_max", C.c_double), ("fr_max", C.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def dptr(a):
    """double* view of a C-contiguous float64 array (None -> NULL)."""
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], "need contiguous float64"
    return a.ctypes.data_as(_d)


def i64ptr(a):
    if a is None:
        return None
    assert a.dtype == np.int64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_i64)


def i32ptr(a):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_i32)


# ---------------------------------------------------------------------------
# declarations (argtypes/restype) shared by both libraries
# ---------------------------------------------------------------------------
class LbfgsbParam(C.Structure):
    """sw2dh_lbfgsb_param (include/sw2d_host.h)."""

    _fields_ = [
        ("m", C.c_int32), ("epsilon", C.c_double), ("epsilon_rel", C.c_double),
        ("past", C.c_int32), ("delta", C.c_double), ("max_iterations", C.c_int32),
        ("max_linesearch", C.c_int32), ("ftol", C.c_double), ("wolfe", C.c_double),
        ("max_step", C.c_double), ("first_step", C.c_int32), ("max_submin", C.c_int32),
        ("restart", C.c_int32),
    ]


# int (*)(const double *x, double *f, double *grad, void *user)
OBJECTIVE = C.CFUNCTYPE(C.c_int, _d, _d, _d, C.c_void_p)

_HOST_DECL = {
    "sw2dh_error": ([], C.c_char_p),
    "sw2dh_lbfgsb_default": ([C.POINTER(LbfgsbParam)], None),
    "sw2dh_lbfgsb": ([C.c_int64, _d, _d, _d, OBJECTIVE, C.c_void_p, C.POINTER(LbfgsbParam), _d, _i32,
                      _i32, _d, _i32], C.c_int),
    "sw2dh_lbfgsb_error": ([], C.c_char_p),
    "sw2dh_geom_rect": ([C.c_int32, C.c_int32, C.c_double, C.c_double, C.POINTER(C.c_void_p)], C.c_int),
    "sw2dh_geom_delaunay": ([C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double, C.c_uint64,
                             C.POINTER(C.c_void_p), _i64], C.c_int),
    "sw2dh_geom_cells": ([C.c_int64, _d, _d, _d, C.c_int64, _i64, _i64, C.POINTER(C.c_void_p)], C.c_int),
    "sw2dh_geom_read": ([C.c_char_p, C.POINTER(C.c_void_p)], C.c_int),
    "sw2dh_geom_write": ([C.c_void_p, C.c_char_p], C.c_int),
    "sw2dh_geom_set_node_z": ([C.c_void_p, _d], C.c_int),
    "sw2dh_geom_set_bnd_codes": ([C.c_void_p, _i32], C.c_int),
    "sw2dh_geom_sizes": ([C.c_void_p, _i64, _i64, _i64, _i64], C.c_int),
    "sw2dh_geom_array": ([C.c_void_p, C.c_char_p, _i64, _i32], C.c_void_p),
    "sw2dh_geom_free": ([C.c_void_p], None),
    "sw2dh_write_case": ([C.c_char_p, C.c_void_p, C.c_char_p, _d, _d, _d, _d, _d, C.c_char_p, _d,
                          C.c_int32, _i32, C.c_char_p, C.c_int32, _d, _d], C.c_int),
    "sw2dh_case_load": ([C.c_char_p, C.POINTER(C.c_void_p)], C.c_int),
    "sw2dh_case_sizes": ([C.c_void_p, _i64, _i64, _i64, _i32, _i32], C.c_int),
    "sw2dh_case_params": ([C.c_void_p, C.POINTER(Params)], C.c_int),
    "sw2dh_case_array": ([C.c_void_p, C.c_char_p, _i64, _i32], C.c_void_p),
    "sw2dh_case_key": ([C.c_void_p, C.c_char_p], C.c_char_p),
    "sw2dh_case_n_decimal": ([C.c_void_p], C.c_int),
    "sw2dh_write_result": ([C.c_void_p, C.c_char_p, _d, _d, _d, C.c_int32, C.c_char_p], C.c_int),
    "sw2dh_result_name": ([C.c_double, C.c_char_p, C.c_int32], C.c_int),
    "sw2dh_case_free": ([C.c_void_p], None),
    "sw2dh_case_probe_name": ([C.c_void_p, C.c_int32], C.c_char_p),
    "sw2dh_geom_locate": ([C.c_void_p, C.c_int64, _d, _d, _i64], C.c_int),
    "sw2dh_case_locate": ([C.c_void_p, C.c_int64, _d, _d, _i64], C.c_int),
}

_GPU_DECL = {
    "sw2dg_version": ([], C.c_char_p),
    "sw2dg_math_eval": ([C.c_int32, C.c_int64, _d, _d, C.c_int32], C.c_int),
    "sw2dg_ctx_create": ([C.POINTER(Mesh), C.POINTER(Params), C.c_int32, C.POINTER(C.c_void_p)], C.c_int),
    "sw2dg_set_cell_params": ([C.c_void_p, _d, _d, _d, _d, _d, _d, _d, _d, _d], C.c_int),
    "sw2dg_set_bc_series": ([C.c_void_p, C.c_int32, C.c_int32, _i32, _d, _d], C.c_int),
    "sw2dg_set_wind": ([C.c_void_p, C.c_int32, C.c_int32, _i32, _d, _d, _d], C.c_int),
    "sw2dg_set_runoff": ([C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, _i32, _d, _d, _d],
                         C.c_int),
    "sw2dg_set_infiltration": ([C.c_void_p, C.c_int32, _d], C.c_int),
    "sw2dg_set_momentum_diffusion": ([C.c_void_p, C.c_int32], C.c_int),
    "sw2dg_set_state": ([C.c_void_p, _d, _d, _d, C.c_double], C.c_int),
    "sw2dg_get_state": ([C.c_void_p, _d, _d, _d], C.c_int),
    "sw2dg_run": ([C.c_void_p, C.c_double, C.c_int64, C.c_int32, _i64], C.c_int),
    "sw2dg_get_clock": ([C.c_void_p, _d, _d, _i64, _i32], C.c_int),
    "sw2dg_get_bnd_f1": ([C.c_void_p, _d], C.c_int),
    "sw2dg_profile_kernels": ([C.c_void_p, C.c_int64, C.c_int32, _d, C.POINTER(C.c_char_p), _i32], C.c_int),
    "sw2dg_device_bytes": ([C.c_void_p], C.c_int64),
    "sw2dg_last_error": ([C.c_void_p], C.c_char_p),
    "sw2dg_ctx_destroy": ([C.c_void_p], None),
}

# every symbol the public headers declare (checked by the CPU test suite)
GPU_SYMBOLS = tuple(_GPU_DECL)
HOST_SYMBOLS = tuple(_HOST_DECL)


def _declare(lib, decl):
    for name, (args, res) in decl.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    return lib


_host = None
_gpu = None


def _load(fname):
    path = os.path.join(LIB_DIR, fname)
    if not os.path.exists(path):
        raise ImportError(
            f"{path} is missing: build it with `make -C {HERE}` (or __graft_entry__.build())")
    return C.CDLL(path)


def host_lib():
    """libsw2d_host.so (no HIP dependency)."""
    global _host
    if _host is None:
        _host = _declare(_load("libsw2d_host.so"), _HOST_DECL)
    return _host


def gpu_lib():
    """libsw2dg.so (HIP kernels).  Raises if it was not built."""
    global _gpu
    if _gpu is None:
        lib = _load("libsw2dg.so")
        _declare(lib, _GPU_DECL)
        _declare(lib, _HOST_DECL)
        _gpu = lib
    return _gpu


def host_check(rc):
    if rc != 0:
        raise SW2DError(rc, host_lib().sw2dh_error().decode())
