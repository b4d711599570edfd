"""ctypes binding of the C ABI declared in include/bld.h.

The library is built in-tree (``libbld.so`` next to this file).  There is no
fallback: every compute call goes through the HIP engine, and a missing
library raises immediately.
"""

import ctypes
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libbld.so")

BLD_ABI_VERSION = 2
BLD_OK, BLD_EARG, BLD_EUNSUPPORTED, BLD_EHIP, BLD_EDIVERGED, BLD_ENOMEM = 0, -1, -2, -3, -4, -5
BLD_FP64, BLD_FP32_MIXED = 0, 1
BLD_BC_DIRICHLET, BLD_BC_NEUMANN = 0, 1
BLD_WALL_FREESLIP, BLD_WALL_ROUGH = 0, 1
BLD_SSPRK33, BLD_SSPRK22, BLD_EULER, BLD_AB2 = 0, 1, 2, 3
# mean-profile kinds (order of bld_profile)
PROFILES = ("vel1", "vel2", "vel3", "adv1", "adv2", "adv3", "sgs11", "sgs12", "sgs13", "sgs22", "sgs23", "sgs33")
I_PROFILES = ("vel3", "adv3", "sgs13", "sgs23")

# kernel classes reported by bld_kernel_times (order of the engine's KernelKind;
# "sgs_viscosity" has no launches: the eddy viscosity is computed inside the
# "sgs_stress" kernel)
KERNEL_KINDS = ("linear", "c2r_x2", "c2r_x1", "adv_products", "sgs_viscosity", "sgs_stress", "r2c_x1",
                "r2c_x2", "stage_project", "check_finite", "sgs_mlp", "c2r_2d", "r2c_2d")

# every symbol include/bld.h declares
EXPORTS = ("bld_create", "bld_destroy", "bld_ensemble_create", "bld_ensemble_size", "bld_select_instance",
           "bld_autotune", "bld_tuning", "bld_last_error", "bld_set_state", "bld_get_state", "bld_rates",
           "bld_project", "bld_init_problem", "bld_step", "bld_step_graph", "bld_set_field_physical",
           "bld_get_field_physical", "bld_time_steps", "bld_set_profiling", "bld_kernel_times", "bld_time_kernel",
           "bld_ldlt_create", "bld_ldlt_destroy", "bld_ldlt_solve", "bld_ldlt_bench", "bld_ldlt_last_error",
           "bld_mlp_nparams", "bld_mlp_create", "bld_mlp_destroy", "bld_mlp_last_error", "bld_mlp_get_params",
           "bld_mlp_forward", "bld_mlp_gradient", "bld_mlp_set_data", "bld_mlp_train", "bld_mlp_time",
           "bld_sgs_closure", "bld_closure_sample", "bld_set_method", "bld_profiles_enable", "bld_profiles_read",
           "bld_set_transform_path", "bld_ldlt_solve_noalloc", "bld_ldlt_noalloc_last_error")


class BldBC(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("reserved", ctypes.c_int32), ("value", ctypes.c_double),
                ("difference", ctypes.c_double)]


class BldConfig(ctypes.Structure):
    _fields_ = [
        ("abi_version", ctypes.c_int32), ("precision", ctypes.c_int32),
        ("k1max", ctypes.c_int32), ("k2max", ctypes.c_int32), ("n3", ctypes.c_int32), ("reserved0", ctypes.c_int32),
        ("L1", ctypes.c_double), ("L2", ctypes.c_double),
        ("alpha_c", ctypes.POINTER(ctypes.c_double)), ("alpha_i", ctypes.POINTER(ctypes.c_double)),
        ("bc", (BldBC * 2) * 3),
        ("diffusion", ctypes.c_int32 * 3), ("reserved1", ctypes.c_int32),
        ("diffusivity", ctypes.c_double * 3),
        ("source", ctypes.c_double * 3),
        ("pressure", ctypes.c_int32),
        ("advection", ctypes.c_int32), ("adv_n1", ctypes.c_int32), ("adv_n2", ctypes.c_int32),
        ("sgs", ctypes.c_int32), ("sgs_n1", ctypes.c_int32), ("sgs_n2", ctypes.c_int32),
        ("sgs_wall", ctypes.c_int32 * 2), ("reserved2", ctypes.c_int32),
        ("sgs_lc", ctypes.POINTER(ctypes.c_double)), ("sgs_li", ctypes.POINTER(ctypes.c_double)),
        ("sgs_factor_strain", ctypes.c_double * 2), ("sgs_factor_stress", ctypes.c_double * 2),
        ("mean", ctypes.c_int32 * 3), ("reserved3", ctypes.c_int32),
        ("mean_value", ctypes.c_double * 3), ("mean_weight", ctypes.POINTER(ctypes.c_double)),
    ]


_lib = None


def load():
    """Load libbld.so and declare the signatures (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"HIP engine library not built: {LIB_PATH} (run __graft_entry__.build())")
    lib = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    D = ctypes.POINTER(ctypes.c_double)
    F = ctypes.POINTER(ctypes.c_float)
    i64 = ctypes.c_int64
    sig = {
        "bld_create": ([ctypes.POINTER(BldConfig), ctypes.POINTER(P)], ctypes.c_int),
        "bld_destroy": ([P], None),
        "bld_ensemble_create": ([ctypes.c_int, ctypes.POINTER(BldConfig), ctypes.POINTER(P)], ctypes.c_int),
        "bld_ensemble_size": ([P], ctypes.c_int),
        "bld_select_instance": ([P, ctypes.c_int], ctypes.c_int),
        "bld_autotune": ([P, ctypes.c_int], ctypes.c_int),
        "bld_tuning": ([P, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int), D], ctypes.c_int),
        "bld_set_transform_path": ([P, ctypes.c_int], ctypes.c_int),
        "bld_ldlt_solve_noalloc": ([i64, i64, P, P, P, ctypes.c_int, P], ctypes.c_int),
        "bld_ldlt_noalloc_last_error": ([], ctypes.c_char_p),
        "bld_last_error": ([P], ctypes.c_char_p),
        "bld_set_state": ([P, D, D, D], ctypes.c_int),
        "bld_get_state": ([P, D, D, D], ctypes.c_int),
        "bld_rates": ([P, D, D, D], ctypes.c_int),
        "bld_project": ([P], ctypes.c_int),
        "bld_init_problem": ([P, ctypes.c_double], ctypes.c_int),
        "bld_step": ([P, ctypes.c_double, i64, D, ctypes.c_double], ctypes.c_int),
        "bld_step_graph": ([P, ctypes.c_double, i64, D, ctypes.c_double], ctypes.c_int),
        "bld_set_field_physical": ([P, ctypes.c_int, D, ctypes.c_int], ctypes.c_int),
        "bld_get_field_physical": ([P, ctypes.c_int, D, ctypes.c_int], ctypes.c_int),
        "bld_time_steps": ([P, ctypes.c_double, i64, ctypes.c_int, D], ctypes.c_int),
        "bld_set_profiling": ([P, ctypes.c_int], ctypes.c_int),
        "bld_kernel_times": ([P, ctypes.c_int, D, ctypes.POINTER(i64)], ctypes.c_int),
        "bld_time_kernel": ([P, ctypes.c_int, ctypes.c_int, D, ctypes.POINTER(i64)], ctypes.c_int),
        "bld_ldlt_create": ([i64, i64, D, D, ctypes.c_int, ctypes.POINTER(P)], ctypes.c_int),
        "bld_ldlt_destroy": ([P], None),
        "bld_ldlt_solve": ([P, D], ctypes.c_int),
        "bld_ldlt_bench": ([P, D, ctypes.c_int, D, D], ctypes.c_int),
        "bld_ldlt_last_error": ([P], ctypes.c_char_p),
        "bld_mlp_nparams": ([], ctypes.c_int),
        "bld_mlp_create": ([F, ctypes.POINTER(P)], ctypes.c_int),
        "bld_mlp_destroy": ([P], None),
        "bld_mlp_last_error": ([P], ctypes.c_char_p),
        "bld_mlp_get_params": ([P, F], ctypes.c_int),
        "bld_mlp_forward": ([P, F, i64, F], ctypes.c_int),
        "bld_mlp_gradient": ([P, F, F, i64, F, D], ctypes.c_int),
        "bld_mlp_set_data": ([P, F, F, i64], ctypes.c_int),
        "bld_mlp_train": ([P, i64, i64, ctypes.c_float, D], ctypes.c_int),
        "bld_mlp_time": ([P, i64, ctypes.c_int, D, D], ctypes.c_int),
        "bld_sgs_closure": ([P, F, ctypes.c_double, D, ctypes.c_int], ctypes.c_int),
        "bld_set_method": ([P, ctypes.c_int], ctypes.c_int),
        "bld_profiles_enable": ([P, ctypes.c_uint32], ctypes.c_int),
        "bld_profiles_read": ([P, ctypes.c_int, D, ctypes.POINTER(i64), ctypes.c_int], ctypes.c_int),
        "bld_closure_sample": ([P, ctypes.c_int, ctypes.c_double, D, F, F, i64, ctypes.POINTER(i64)], ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


class BldError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


def dptr(a):
    """double* of a contiguous float64/complex128 array (or NULL for None)."""
    if a is None:
        return ctypes.cast(None, ctypes.POINTER(ctypes.c_double))
    assert a.dtype in (np.float64, np.complex128) and (a.flags.f_contiguous or a.flags.c_contiguous)
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def fptr(a):
    """float* of a contiguous float32 array."""
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
