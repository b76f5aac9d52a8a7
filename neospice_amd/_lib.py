"""ctypes binding of libneospice_amd.so (the C ABI in include/neospice_amd.h).

The library must have been built (neospice_amd/build.py); importing this
module fails loudly otherwise -- there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libneospice_amd.so")

I32 = C.c_int32
P_I32 = C.POINTER(C.c_int32)
P_I64 = C.POINTER(C.c_int64)
P_F64 = C.POINTER(C.c_double)
VP = C.c_void_p


class Options(C.Structure):
    _fields_ = [(k, C.c_double) for k in
                ("abstol", "reltol", "vntol", "trtol", "chgtol", "gmin", "diag_gmin", "xmu", "src_fact")] + \
               [(k, C.c_int32) for k in ("itl1", "itl4", "method", "engine")]


class Ctx(C.Structure):
    _fields_ = [("mode", C.c_int32), ("order", C.c_int32), ("method", C.c_int32), ("pad", C.c_int32),
                ("ag0", C.c_double), ("ag1", C.c_double), ("ag2", C.c_double),
                ("xmu_ratio", C.c_double), ("time", C.c_double), ("src_fact", C.c_double),
                ("delta", C.c_double), ("delta_old1", C.c_double), ("delta_old2", C.c_double)]


_SIGS = {
    "nsa_last_error": (C.c_char_p, []),
    "nsa_default_options": (None, [C.POINTER(Options)]),
    "nsa_circuit_create": (VP, []),
    "nsa_circuit_destroy": (None, [VP]),
    "nsa_circuit_node": (C.c_int, [VP, C.c_char_p, P_I32]),
    "nsa_circuit_r": (C.c_int, [VP, C.c_char_p, I32, I32, C.c_double, P_I32]),
    "nsa_circuit_c": (C.c_int, [VP, C.c_char_p, I32, I32, C.c_double, P_I32]),
    "nsa_circuit_v": (C.c_int, [VP, C.c_char_p, I32, I32, C.c_double, P_I32]),
    "nsa_circuit_v_pulse": (C.c_int, [VP, C.c_char_p, I32, I32, P_F64, P_I32]),
    "nsa_circuit_v_pwl": (C.c_int, [VP, C.c_char_p, I32, I32, I32, P_F64, P_F64, P_I32]),
    "nsa_circuit_i": (C.c_int, [VP, C.c_char_p, I32, I32, C.c_double, P_I32]),
    "nsa_circuit_model": (C.c_int, [VP, C.c_char_p]),
    "nsa_circuit_d": (C.c_int, [VP, C.c_char_p, I32, I32, C.c_char_p, P_I32]),
    "nsa_circuit_m": (C.c_int, [VP, C.c_char_p, I32, I32, I32, I32, C.c_char_p, C.c_double,
                                C.c_double, P_I32]),
    "nsa_circuit_bsim4_params": (C.c_int, [VP, I32, P_F64]),
    "nsa_circuit_bsim4_param_derivs": (C.c_int, [VP, I32, P_F64, P_F64, P_F64]),
    "nsa_circuit_finalize": (C.c_int, [VP]),
    "nsa_circuit_info": (C.c_int, [VP, P_I32]),
    "nsa_circuit_pattern": (C.c_int, [VP, P_I32, P_I32]),
    "nsa_circuit_load_order": (C.c_int, [VP, P_I32]),
    "nsa_circuit_branch": (C.c_int, [VP, I32, P_I32]),
    "nsa_engine_create": (C.c_int, [VP, I32, C.POINTER(Options), C.POINTER(VP)]),
    "nsa_engine_destroy": (None, [VP]),
    "nsa_engine_set_values": (C.c_int, [VP, I32, P_F64]),
    "nsa_engine_set_model_values": (C.c_int, [VP, I32, C.c_char_p, P_F64]),
    "nsa_engine_set_ctx": (C.c_int, [VP, C.POINTER(Ctx), I32]),
    "nsa_engine_newton": (C.c_int, [VP, I32, P_I32, P_I32]),
    "nsa_engine_get_solution": (C.c_int, [VP, P_F64]),
    "nsa_engine_set_solution": (C.c_int, [VP, P_F64]),
    "nsa_engine_get_state": (C.c_int, [VP, I32, P_F64]),
    "nsa_engine_set_state": (C.c_int, [VP, I32, P_F64]),
    "nsa_engine_seed_states": (C.c_int, [VP]),
    "nsa_engine_rotate_states": (C.c_int, [VP]),
    "nsa_engine_trunc": (C.c_int, [VP, P_F64, P_F64]),
    "nsa_engine_trunc_devices": (C.c_int, [VP, P_I32]),
    "nsa_engine_reset_solver": (C.c_int, [VP]),
    "nsa_engine_load": (C.c_int, [VP, P_F64, P_F64]),
    "nsa_engine_solve_transposed": (C.c_int, [VP, P_F64]),
    "nsa_engine_family_count": (C.c_int, [VP, I32, P_I32]),
    "nsa_engine_set_family_values": (C.c_int, [VP, I32, P_F64]),
    "nsa_engine_get_family_values": (C.c_int, [VP, I32, P_F64]),
    "nsa_engine_dc_adjoint": (C.c_int, [VP, P_F64, P_F64, P_F64, P_F64, P_F64, P_F64]),
    "nsa_engine_tran_forward": (C.c_int, [VP, I32, C.c_double, P_I32, I32, I32, P_F64]),
    "nsa_engine_tran_adjoint": (C.c_int, [VP, P_F64, P_F64, P_F64]),
    "nsa_engine_adjoint_timing": (C.c_int, [VP, P_F64]),
    "nsa_engine_solver_name": (C.c_char_p, [VP]),
    "nsa_engine_stats": (C.c_int, [VP, P_I64]),
    "nsa_engine_set_timing": (C.c_int, [VP, I32]),
    "nsa_engine_timing": (C.c_int, [VP, P_F64]),
    "nsa_lu_create": (C.c_int, [I32, I32, P_I32, P_I32, I32, I32, C.POINTER(VP)]),
    "nsa_lu_destroy": (None, [VP]),
    "nsa_lu_numeric": (C.c_int, [VP, P_F64, C.c_double, P_I32]),
    "nsa_lu_refactorize": (C.c_int, [VP, P_F64, C.c_double, P_I32]),
    "nsa_lu_set_equilibrate": (C.c_int, [VP, I32]),
    "nsa_lu_numeric_complex": (C.c_int, [VP, P_F64, P_I32]),
    "nsa_lu_refactorize_complex": (C.c_int, [VP, P_F64, P_I32]),
    "nsa_lu_solve_complex": (C.c_int, [VP, P_F64]),
    "nsa_lu_ac_sweep": (C.c_int, [VP, P_F64, P_F64, P_F64, I32, P_F64, P_F64, P_I32]),
    "nsa_lu_ac_stats": (C.c_int, [VP, P_F64]),
    "nsa_bundle_set_tuning": (C.c_int, [VP, I32]),
    "nsa_bundle_tuning": (C.c_int, [VP, P_F64, I32]),
    "nsa_lu_solve": (C.c_int, [VP, P_F64]),
    "nsa_lu_solve_transposed": (C.c_int, [VP, P_F64]),
    "nsa_lu_solve_noalloc": (C.c_int, [VP, P_F64]),
    "nsa_lu_name": (C.c_char_p, [VP]),
    "nsa_lu_info": (C.c_int, [VP, P_I32]),
    "nsa_lu_structure": (C.c_int, [VP, P_I32, P_I32, P_I32, P_I32]),
    "nsa_pivot_structure": (C.c_int, [I32, I32, P_I32, P_I32, P_F64, C.c_double, P_I32, P_I32, P_I32,
                                      P_I32, P_I32, P_I32]),
    "nsa_min_degree_order": (C.c_int, [I32, P_I32, P_I32, P_I32]),
    "nsa_bundle_create": (C.c_int, [C.POINTER(VP), I32, C.POINTER(Options), C.POINTER(VP)]),
    "nsa_bundle_destroy": (None, [VP]),
    "nsa_bundle_info": (C.c_int, [VP, P_I32]),
    "nsa_bundle_offsets": (C.c_int, [VP, P_I32, P_I32]),
    "nsa_bundle_dc": (C.c_int, [VP, I32, P_I32, P_I32]),
    "nsa_bundle_get_solution": (C.c_int, [VP, P_F64]),
    "nsa_bundle_set_solution": (C.c_int, [VP, P_F64]),
    "nsa_bundle_tran_begin": (C.c_int, [VP, C.c_double, C.c_double, P_I32, P_I32, I32]),
    "nsa_bundle_tran_advance": (C.c_int, [VP, C.c_double, P_I32]),
    "nsa_bundle_tran_status": (C.c_int, [VP, I32, P_I64]),
    "nsa_bundle_tran_steps": (C.c_int, [VP, I32, P_F64, P_F64, P_I32, P_I32, P_I32]),
    "nsa_bundle_tran_probes": (C.c_int, [VP, I32, P_I32, P_F64]),
    "nsa_bundle_stats": (C.c_int, [VP, P_I64]),
    "nsa_bundle_set_timing": (C.c_int, [VP, I32]),
    "nsa_bundle_timing": (C.c_int, [VP, P_F64]),
    "nsa_bundle_set_graph": (C.c_int, [VP, I32]),
    "nsa_bundle_set_device_control": (C.c_int, [VP, I32]),
    "nsa_bundle_tick_stats": (C.c_int, [VP, P_F64]),
    "nsa_bundle_plan_info": (C.c_int, [VP, I32, P_I64]),
    "nsa_bundle_device_offsets": (C.c_int, [VP, P_I32]),
    "nsa_bundle_dc_adjoint": (C.c_int, [VP, P_F64, P_F64, P_F64, P_F64, P_F64]),
    "nsa_bundle_set_device_values": (C.c_int, [VP, P_F64, P_F64, P_F64]),
    "nsa_bundle_get_device_values": (C.c_int, [VP, P_F64, P_F64, P_F64]),
    "nsa_bundle_dc_ladder": (C.c_int, [VP, I32, P_I32, P_I32, P_I32]),
}

_lib = None


def lib():
    """Loads the library once.  Raises if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("neospice_amd: %s is missing; run neospice_amd/build.py" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class NsaError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("neospice_amd error %d: %s" % (code, msg))
        self.code = code


def check(rc):
    if rc < 0:
        raise NsaError(rc, lib().nsa_last_error().decode())
    return rc


def f64p(a):
    return a.ctypes.data_as(P_F64)


def i32p(a):
    return a.ctypes.data_as(P_I32)


def as_f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def as_i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)
