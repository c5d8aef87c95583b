"""ctypes binding of the C-ABI in include/fsilbm3d_amd.h (libfsil.so).

The library is the only compute path: there is no CPU fallback, and loading
fails loudly when the shared object has not been built.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfsil.so")

# every exported symbol with (restype, argtypes); the CPU test checks this
# list against include/fsilbm3d_amd.h
_i, _d, _pi, _pd = C.c_int, C.c_double, C.POINTER(C.c_int), C.POINTER(C.c_double)
_pf = C.POINTER(C.c_float)
SIGNATURES = {
    "fsil_version": (_i, []),
    "fsil_last_error": (C.c_char_p, []),
    "fsil_init": (_i, [_i]),
    "fsil_finalize": (_i, []),
    "fsil_synchronize": (_i, []),
    "fsil_block_create": (_i, [_i, _pi, _pd, _pi, _i, _pd, _pd, _pd, _i, _pi]),
    "fsil_block_destroy": (_i, [_i]),
    "fsil_set_flow": (_i, [_i, _d, _pd, _pd, _i, _pd, _d, _d, _d]),
    "fsil_set_time": (_i, [_i, _d]),
    "fsil_get_volume_force": (_i, [_i, _pd]),
    "fsil_f_upload": (_i, [_i, _pd]),
    "fsil_f_fill": (_i, [_i, _pd]),
    "fsil_f_download": (_i, [_i, _pd]),
    "fsil_set_boundary_conditions": (_i, [_i]),
    "fsil_macro_download": (_i, [_i, _pd, _pd]),
    "fsil_collide_stream_bc": (_i, [_i]),
    "fsil_steps": (_i, [_i, _i, _d, _d]),
    "fsil_time_collide": (_i, [_i, _i, _pd]),
    "fsil_time_steps": (_i, [_i, _i, _d, _d, _pd]),
    "fsil_ibm_define_bodies": (_i, [_i, _i, _pi, _pi, _pi, _pi, _pi]),
    "fsil_ibm_set_params": (_i, [_i, _i, _d, _d, _d]),
    "fsil_ibm_interaction_force": (_i, [_i, _d, _pd, _pd, _pd, _pd, _pd, _pd, _pi, _pd]),
    "fsil_ibm_band_info": (_i, [_i, _pi, _pi]),
    "fsil_ibm_forces": (_i, [_i, _pd, _pd]),
    "fsil_ibm_report": (_i, [_i, _pi, _pd]),
    "fsil_copy_bandwidth": (_i, [C.c_longlong, _i, _pd]),
    "fsil_tree_add_son": (_i, [_i, _i, _i, _pi]),
    "fsil_tree_pair_info": (_i, [_i, _i, _pi]),
    "fsil_extract_interpolate_layer": (_i, [_i, _i]),
    "fsil_interpolation_father_to_son": (_i, [_i, _i, _i]),
    "fsil_deliver_son_to_father": (_i, [_i, _i]),
    "fsil_tree_steps": (_i, [_i, _i, _d, _d]),
    "fsil_time_tree_steps": (_i, [_i, _i, _d, _d, _pd]),
    "fsil_mlp_create": (_i, [_i, _pi]),
    "fsil_mlp_destroy": (_i, [_i]),
    "fsil_mlp_get_params": (_i, [_i, _pf]),
    "fsil_mlp_set_params": (_i, [_i, _pf]),
    "fsil_mlp_forward": (_i, [_i, _pf, _i, _pf, _i]),
    "fsil_mlp_loss_grad": (_i, [_i, _pf, _pf, _i, _pd, _pf]),
    "fsil_mlp_train": (_i, [_i, _pf, _pf, _i, _i, _i, _d, _i, _pd, _pd]),
    "fsil_ibm_set_warm_start": (_i, [_i, _i, _i, _i]),
    "fsil_ibm_rounds": (_i, [_i, _pi]),
    "fsil_fsi_step": (_i, [_i, _d, _i, _pi, _pd]),
    "fsil_set_viscosity": (_i, [_i, _d]),
    "fsil_flow_pack": (_i, [_i, _i, _d, _d, _pf]),
    "fsil_write_flow": (_i, [_i, _i, _d, _d, C.c_char_p]),
    "fsil_write_wait": (_i, []),
    "fsil_write_continue": (_i, [_i, _pi, _i, _d, C.c_char_p]),
    "fsil_fem_checkpoint": (_i, [_i, C.c_void_p, C.POINTER(C.c_longlong)]),
    "fsil_fem_restore": (_i, [_i, C.c_void_p, C.c_longlong]),
    "fsil_tau_download": (_i, [_i, _pd]),
    "fsil_solid_info": (_i, [_i, _i, _pd, _pd]),
    "fsil_solid_info_static": (_i, [_i, _i, _pd, _pd]),
    "fsil_ibm_warm_start_data": (_i, [_i, _pf, _pf, _pi]),
    "fsil_fem_define_bodies": (_i, [_i, _i, _pi, _pi, _pd, _pi, _pi, _pi, _pi, _pd, _pd, _d, _d, _d, _d]),
    "fsil_fem_set_solver": (_i, [_i, _d, _d, _d, _d, _d, _d, _i]),
    "fsil_fem_update_markers": (_i, [_i]),
    "fsil_fem_substeps": (_i, [_i, _d, _i, _d]),
    "fsil_fem_state": (_i, [_i, _i, _pd, _pd, _pd, _pd]),
    "fsil_fem_info": (_i, [_i, _i, _pi, _pi, _pd]),
}

_lib = None
_lock = threading.Lock()


class FsilError(RuntimeError):
    pass


def load():
    """Load libfsil.so (built by fsilbm3d_amd/build.py); raise if absent."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise FsilError(f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
                            " (there is no CPU fallback)")
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return lib


def check(rc: int):
    if rc != 0:
        raise FsilError(load().fsil_last_error().decode())


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], (a.dtype, a.flags)
    return a.ctypes.data_as(_pd)


def iptr(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"], (a.dtype, a.flags)
    return a.ctypes.data_as(_pi)


def fptr(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"], (a.dtype, a.flags)
    return a.ctypes.data_as(_pf)


_initialised = False


def init(device: int = 0):
    global _initialised
    lib = load()
    if not _initialised:
        check(lib.fsil_init(device))
        _initialised = True
    return lib
