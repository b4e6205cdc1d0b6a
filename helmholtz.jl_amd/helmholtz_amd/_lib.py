"""ctypes binding of libhelmholtz_amd.so (declarations in include/helmholtz_amd.h).

Loading fails loudly: there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libhelmholtz_amd.so")

HH_OK, HH_ERR_ARG, HH_ERR_HIP, HH_ERR_NOMEM, HH_ERR_UNSUPPORTED = 0, 1, 2, 3, 4
HH_C128, HH_C64 = 0, 1
HH_HOST, HH_DEVICE = 0, 1
HH_FGMRES, HH_BICGSTAB = 0, 1
HH_CONVERGED, HH_MAXITER, HH_NONFINITE, HH_ZERO_RHS = 0, 1, 2, 3

# every symbol include/helmholtz_amd.h declares
EXPORTS = (
    "hh_last_error", "hh_version", "hh_abl", "hh_op_create", "hh_op_destroy", "hh_op_apply",
    "hh_solver_create", "hh_solver_copy", "hh_solver_clear", "hh_solver_destroy", "hh_solve",
    "hh_solver_levels", "hh_device_alloc", "hh_device_free", "hh_memcpy", "hh_op_time_kernel",
    "hh_unet_num_params", "hh_unet_create", "hh_unet_destroy", "hh_unet_get_params", "hh_unet_set_params",
    "hh_unet_forward", "hh_unet_loss_grad", "hh_unet_adam_step", "hh_dataset_create", "hh_dataset_destroy",
    "hh_dataset_size", "hh_solver_capture_begin", "hh_solver_capture_count", "hh_dataset_add_from_solver",
    "hh_unet_train", "hh_unet_eval", "hh_solver_attach_unet", "hh_stream_solve", "hh_solver_precondition",
    "hh_solver_time_kernel", "hh_solver_kernel_time", "hh_op_create_ho", "hh_sommerfeld_bc",
    "hh_abl_faces", "hh_solver_learned_policy", "hh_solver_learned_workspace", "hh_op_apply_noalloc",
)
HH_LEARN_TRIAL, HH_LEARN_ALWAYS = 0, 1
HH_RELAX_JAC, HH_RELAX_JAC_GMRES = 0, 1
HH_KT_NONE, HH_KT_DOWN, HH_KT_MID, HH_KT_UP, HH_KT_APPLY, HH_KT_DOTS, HH_KT_UPDATE = range(7)


class HelmholtzError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[helmholtz_amd error {code}] {msg}")
        self.code = code


class MGParams(C.Structure):
    _fields_ = [
        ("levels", C.c_int32), ("pre", C.c_int32), ("post", C.c_int32), ("relax", C.c_double),
        ("cycle", C.c_int32), ("max_outer", C.c_int32), ("tol", C.c_double), ("inner", C.c_int32),
        ("krylov", C.c_int32), ("prec", C.c_int32), ("learned", C.c_int32), ("verbose", C.c_int32),
        ("graph", C.c_int32), ("autotune", C.c_int32), ("fused", C.c_int32), ("wlevels", C.c_int32),
        ("per_level", C.c_int32), ("pre_l", C.c_int32 * 16), ("post_l", C.c_int32 * 16),
        ("relax_type", C.c_int32), ("basis", C.c_int32),
    ]


class SolveStats(C.Structure):
    _fields_ = [
        ("iters", C.POINTER(C.c_int32)), ("status", C.POINTER(C.c_int32)),
        ("true_res", C.POINTER(C.c_double)), ("max_iters", C.c_int32), ("n_converged", C.c_int32),
        ("setup_time", C.c_double), ("solve_time", C.c_double), ("n_prec", C.c_int64),
        ("learned_used", C.c_int32), ("fell_back", C.c_int32),
    ]


_lib = None


def lib():
    """Load the library once; raise if it is missing (build it with build.py)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise HelmholtzError(-1, f"{LIB_PATH} not built; run helmholtz.jl_amd/build.py")
    L = C.CDLL(LIB_PATH)
    vp, i32, i64, dbl = C.c_void_p, C.c_int, C.c_int64, C.c_double
    p64 = C.POINTER(C.c_int64)
    pd = C.POINTER(C.c_double)
    L.hh_last_error.restype = C.c_char_p
    L.hh_last_error.argtypes = []
    L.hh_version.restype = i32
    L.hh_abl.argtypes = [i32, p64, i32, p64, dbl, vp, i32]
    L.hh_abl_faces.argtypes = [i32, p64, i32, p64, dbl, C.POINTER(C.c_int32), vp, i32]
    L.hh_sommerfeld_bc.argtypes = [i32, p64, pd, vp, dbl, i32, i32, vp, i32]
    L.hh_op_create.argtypes = [i32, p64, pd, vp, vp, dbl, dbl, i32, i32, i32, C.POINTER(vp)]
    L.hh_op_create_ho.argtypes = [i32, p64, pd, vp, vp, dbl, dbl, i32, i32, dbl, dbl, C.POINTER(vp)]
    L.hh_op_destroy.argtypes = [vp]
    L.hh_op_apply.argtypes = [vp, dbl, vp, vp, i64, i32, i32]
    L.hh_solver_create.argtypes = [vp, C.POINTER(MGParams), pd, C.POINTER(vp)]
    L.hh_solver_copy.argtypes = [vp, C.POINTER(vp)]
    L.hh_solver_clear.argtypes = [vp]
    L.hh_solver_destroy.argtypes = [vp]
    L.hh_solve.argtypes = [vp, vp, vp, i64, i32, i32, C.POINTER(SolveStats)]
    L.hh_solver_levels.argtypes = [vp, C.POINTER(C.c_int), vp, i32]
    L.hh_device_alloc.argtypes = [i64, C.POINTER(vp)]
    L.hh_device_free.argtypes = [vp]
    L.hh_memcpy.argtypes = [vp, vp, i64, i32]
    L.hh_op_time_kernel.argtypes = [vp, i32, dbl, vp, vp, vp, i64, i32, i32, pd]
    f32, u64 = C.c_float, C.c_uint64
    pf = C.POINTER(C.c_float)
    L.hh_unet_num_params.argtypes = []
    L.hh_unet_create.argtypes = [u64, C.POINTER(vp)]
    L.hh_unet_destroy.argtypes = [vp]
    L.hh_unet_get_params.argtypes = [vp, vp]
    L.hh_unet_set_params.argtypes = [vp, vp]
    L.hh_unet_forward.argtypes = [vp, i32, i32, i32, vp, vp]
    L.hh_unet_loss_grad.argtypes = [vp, i32, i32, i32, vp, vp, pf, vp]
    L.hh_unet_adam_step.argtypes = [vp, vp, f32, f32, f32, f32]
    L.hh_dataset_create.argtypes = [i32, i32, i64, C.POINTER(vp)]
    L.hh_dataset_destroy.argtypes = [vp]
    L.hh_dataset_size.argtypes = [vp, C.POINTER(C.c_int64)]
    L.hh_solver_capture_begin.argtypes = [vp, i32, i64, i32, i32]
    L.hh_solver_capture_count.argtypes = [vp, C.POINTER(C.c_int64)]
    L.hh_dataset_add_from_solver.argtypes = [vp, vp, i32]
    L.hh_unet_train.argtypes = [vp, vp, i32, i32, f32, u64, pf]
    L.hh_unet_eval.argtypes = [vp, vp, i64, i64, pf]
    L.hh_solver_attach_unet.argtypes = [vp, vp, i32]
    L.hh_solver_learned_policy.argtypes = [vp, i32]
    L.hh_op_apply_noalloc.argtypes = [vp, C.c_double, vp, vp, i64, i32]
    L.hh_solver_learned_workspace.argtypes = [vp, C.POINTER(vp), C.POINTER(vp)]
    L.hh_solver_precondition.argtypes = [vp, vp, vp, i64, i32]
    L.hh_solver_time_kernel.argtypes = [vp, i32, i32]
    L.hh_solver_kernel_time.argtypes = [vp, pd, p64, pd]
    for name in EXPORTS:
        if name == "hh_last_error":
            continue
        getattr(L, name).restype = C.c_int64 if name == "hh_unet_num_params" else C.c_int
    _lib = L
    return L


def check(code: int):
    if code != HH_OK:
        raise HelmholtzError(code, lib().hh_last_error().decode())
