"""ctypes binding of libamaru_hip.so (the C ABI in include/amaru_hip.h).

There is no fallback: if the library is missing or does not load, every
entry point raises.  The signatures below are the ones a Julia ``ccall``
glue would use.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libamaru_hip.so")

AMR_OK = 0
AMR_HEX8, AMR_HEX20, AMR_TET10 = 1, 2, 3
AMR_MAT_LE, AMR_MAT_VM, AMR_MAT_DP = 0, 1, 2
AMR_PREC_FP64, AMR_PREC_MIXED, AMR_PREC_MIXED_F64ACC = 0, 1, 2
AMR_PRECOND_JACOBI, AMR_PRECOND_BLOCK_JACOBI, AMR_PRECOND_TWO_LEVEL = 0, 1, 2
ERRORS = {-1: "AMR_ERR_ARG", -2: "AMR_ERR_HIP", -3: "AMR_ERR_JACOBIAN", -4: "AMR_ERR_MATERIAL",
          -5: "AMR_ERR_SOLVE", -6: "AMR_ERR_NAN", -7: "AMR_ERR_NOMEM"}

# every function the header declares, with (restype, argtypes)
P = ctypes.c_void_p
I32, I64, D = ctypes.c_int, ctypes.c_int64, ctypes.c_double


class SolveOpts(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int), ("precond", ctypes.c_int), ("maxiter", ctypes.c_int),
                ("check_every", ctypes.c_int), ("rtol", ctypes.c_double), ("inner_rtol", ctypes.c_double),
                ("replace_ratio", ctypes.c_double), ("drift_limit", ctypes.c_double)]


class SolveInfo(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("refinements", ctypes.c_int), ("converged", ctypes.c_int),
                ("fallback", ctypes.c_int), ("relres", ctypes.c_double)]


class NRInfo(ctypes.Structure):
    _fields_ = [("res", ctypes.c_double), ("err", ctypes.c_double), ("relres", ctypes.c_double),
                ("pcg_iterations", ctypes.c_int), ("status", ctypes.c_int), ("replacements", ctypes.c_int),
                ("fallback", ctypes.c_int)]


SIGNATURES = {
    "amr_last_error": (ctypes.c_char_p, []),
    "amr_device_count": (I32, [P]),
    "amr_model_create": (I32, [P, I32, I32, I32, I64, P, I64, P, I32, P, P, P, P, I64, I64]),
    "amr_model_destroy": (I32, [P]),
    "amr_model_set_dofs": (I32, [P, P, I64, I64]),
    "amr_model_info": (I32, [P, P]),
    "amr_state_set": (I32, [P, P]),
    "amr_state_get": (I32, [P, P, I32]),
    "amr_state_commit": (I32, [P]),
    "amr_state_restore": (I32, [P]),
    "amr_assemble": (I32, [P]),
    "amr_get_Ke": (I32, [P, P]),
    "amr_get_K_csc": (I32, [P, P, P, P, P]),
    "amr_solve_system": (I32, [P, P, P, ctypes.POINTER(SolveOpts), ctypes.POINTER(SolveInfo)]),
    "amr_update_state": (I32, [P, P, P]),
    "amr_stage_begin": (I32, [P, P, P]),
    "amr_increment_begin": (I32, [P, D, D, I32]),
    "amr_nr_iteration": (I32, [P, I32, ctypes.POINTER(SolveOpts), ctypes.POINTER(NRInfo)]),
    "amr_increment_commit": (I32, [P]),
    "amr_increment_reject": (I32, [P]),
    "amr_stage_get": (I32, [P, P, P, I32]),
    "amr_timing": (I32, [P, I32, P, P]),
    "amr_set_spmv_lanes": (I32, [P, I32, I32]),
    "amr_get_spmv_lanes": (I32, [P, P]),
    "amr_tune_spmv_lanes": (I32, [P]),
    "amr_nr_prepare": (I32, [P, I32]),
    "amr_nr_finish": (I32, [P, ctypes.POINTER(SolveInfo), ctypes.POINTER(NRInfo)]),
    "amr_nr_finish_begin": (I32, [P, ctypes.POINTER(SolveInfo)]),
    "amr_nr_prepare_many": (I32, [P, P, I32]),
    "amr_nr_finish_begin_many": (I32, [P, P, I32]),
    "amr_nr_finish_end": (I32, [P, ctypes.POINTER(NRInfo)]),
    "amr_pool_create": (I32, [P, I32, I32, I32]),
    "amr_pool_destroy": (I32, [P]),
    "amr_pool_attach": (I32, [P, P]),
    "amr_pool_detach": (I32, [P]),
    "amr_pool_add": (I32, [P, P, ctypes.POINTER(SolveOpts)]),
    "amr_pool_step": (I32, [P, I32, P, P, ctypes.POINTER(SolveInfo)]),
    "amr_pool_step_begin": (I32, [P]),
    "amr_pool_step_done": (I32, [P, P]),
    "amr_pool_release_left": (I32, [P]),
    "amr_pool_step_end": (I32, [P, I32, P, P, ctypes.POINTER(SolveInfo)]),
    "amr_pool_size": (I32, [P, P]),
    "amr_pool_timing": (I32, [P, I32, P, P]),
    "amr_pool_set_graphs": (I32, [P, I32, P, P]),
    "amr_pool_stats": (I32, [P, P]),
    "amr_model_device_bytes": (I32, [P, P]),
    "amr_mem_info": (I32, [I32, P, P]),
    "amr_mem_trim": (I32, [I32]),
    "amr_set_coarse_refresh": (I32, [P, I32]),
    "amr_set_coarse": (I32, [P, I32, I32]),
    "amr_get_coarse": (I32, [P, P]),
    "amr_get_coarse_space": (I32, [P, P, P]),
    "amr_check_expr": (I32, [P, I32, I32]),
    "amr_cell_loads": (I32, [I32, ctypes.c_int64, I32, I32, I32, P, P, P, P, I32, P, P, I32, I32, D, P]),
    "amr_block_mesh": (I32, [I32, I32, P, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, P, P, P, P]),
    "amr_dof_map": (I32, [I32, ctypes.c_int64, P, P, P, P, P]),
    "amr_dof_map_noalloc": (I32, [I32, ctypes.c_int64, P, P, P, P, P]),
    "amr_assemble_M": (I32, [P, P]),
    "amr_combine_KM": (I32, [P, D, D]),
    "amr_spmv_col_bytes": (I32, [P, P]),
    "amr_set_spmv_col_bytes": (I32, [P, I32]),
    "amr_set_scheme": (I32, [P, I32]),
    "amr_set_elem_tiles": (I32, [P, I32, I32]),
    "amr_set_ke_path": (I32, [P, I32]),
    "amr_get_elem_tiles": (I32, [P, P, P]),
    "amr_tune_elem_tiles": (I32, [P, P, P]),
    "amr_matvec": (I32, [P, I32, P, P]),
    "amr_get_M_csc": (I32, [P, P, P, P, P]),
    "amr_ip_values": (I32, [P, I32, P]),
    "amr_nodal_recovery": (I32, [P, I32, P, P]),
}

AMR_NFIELDS = 19

_lib = None


class AmaruError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code
        self.msg = msg


def lib():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with amaru.jl_amd/build.py "
                           "(the HIP path has no CPU fallback)")
    L = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(rc):
    if rc != AMR_OK:
        raise AmaruError(rc, lib().amr_last_error().decode(errors="replace"))
    return rc


def ptr(a):
    if a is None:
        return None
    assert isinstance(a, np.ndarray) and a.flags["C_CONTIGUOUS"], "need a C-contiguous array"
    return a.ctypes.data_as(ctypes.c_void_p)


def device_count():
    n = ctypes.c_int(0)
    check(lib().amr_device_count(ctypes.byref(n)))
    return n.value
