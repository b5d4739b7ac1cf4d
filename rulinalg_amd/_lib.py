"""ctypes binding of librlgpu.so (the C ABI declared in include/rlgpu.h).

The shared library is built in-tree by `make -C rulinalg_amd` (or
`__graft_entry__.build()`).  Loading fails loudly when it is missing: there is
no CPU fallback behind these calls.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librlgpu.so")

RLG_OK = 0
STATUS_NAMES = {
    0: "Ok", 1: "InvalidArg", 2: "DecompFailure", 3: "AlgebraFailure", 4: "DivByZero",
    5: "ScalarConversionFailure", 6: "InvalidPermutation", 7: "Shape", 8: "Device", 9: "OutOfMemory",
    10: "NotConverged",
}

PREC_F64, PREC_F32, PREC_F16, PREC_BF16, PREC_F16X3 = 0, 1, 2, 3, 4
REFINE_IR, REFINE_GMRES = 0, 1
PIVOT_PARTIAL, PIVOT_TOURNAMENT, PIVOT_AUTO = 0, 1, 2
KIND_LU, KIND_CHOL = 0, 1

sz = C.c_size_t
i32 = C.c_int
i64 = C.c_int64
dbl = C.c_double
flt = C.c_float
vp = C.c_void_p
PD = C.POINTER(C.c_double)
PF = C.POINTER(C.c_float)
PSZ = C.POINTER(C.c_size_t)
PI64 = C.POINTER(C.c_int64)


class MixedOpts(C.Structure):
    _fields_ = [("factor_prec", i32), ("refine", i32), ("tol", dbl), ("max_iter", i32),
                ("fallback", i32), ("pivot", i32)]


class MixedReport(C.Structure):
    _fields_ = [("iterations", i32), ("inner_iterations", i32), ("fell_back", i32),
                ("converged", i32), ("backward_error", dbl), ("factor_ms", dbl),
                ("refine_ms", dbl), ("fallback_ms", dbl)]


class BatchItem(C.Structure):
    _fields_ = [("kind", i32), ("n", i32), ("nrhs", i32), ("status", i32), ("a", vp), ("b", vp),
                ("ipiv", vp)]


class ServerOpts(C.Structure):
    _fields_ = [("mem_budget_bytes", sz), ("slots", i32), ("mixed", i32), ("tol", dbl)]


class Request(C.Structure):
    _fields_ = [("kind", i32), ("n", i32), ("nrhs", i32), ("a", PD), ("b", PD), ("factor_id", i64)]


class Result(C.Structure):
    _fields_ = [("ticket", i64), ("status", i32), ("fell_back", i32), ("backward_error", dbl),
                ("latency_ms", dbl), ("x", PD)]


# name -> (restype, argtypes)
SIGNATURES = {
    "rlg_version": (C.c_char_p, []),
    "rlg_last_error_message": (C.c_char_p, []),
    "rlg_set_device": (i32, [i32]),
    "rlg_synchronize": (i32, []),
    "rlg_dev_alloc": (i32, [sz, C.POINTER(vp)]),
    "rlg_dev_free": (i32, [vp]),
    "rlg_dev_upload": (i32, [vp, vp, sz]),
    "rlg_dev_download": (i32, [vp, vp, sz]),
    "rlg_dev_copy": (i32, [vp, vp, sz, vp]),
    "rlg_gemm_f64": (i32, [sz, sz, vp, sz, sz, sz, vp, sz, vp, sz]),
    "rlg_gemm_f32": (i32, [sz, sz, vp, sz, sz, sz, vp, sz, vp, sz]),
    "rlg_gemm_f64_dev": (i32, [sz, sz, sz, dbl, vp, sz, vp, sz, dbl, vp, sz, vp]),
    "rlg_gemm_f32_dev": (i32, [sz, sz, sz, flt, vp, sz, vp, sz, flt, vp, sz, vp]),
    "rlg_gemm_f32_split_dev": (i32, [sz, sz, sz, flt, vp, sz, vp, sz, flt, vp, sz, vp]),
    "rlg_matvec_f64": (i32, [sz, sz, vp, sz, sz, vp, vp]),
    "rlg_solve_u_triangular_f64": (i32, [sz, sz, vp, sz, sz, sz, vp, sz, vp, sz]),
    "rlg_solve_l_triangular_f64": (i32, [sz, sz, vp, sz, sz, sz, vp, sz, vp, sz]),
    "rlg_solve_u_triangular_f32": (i32, [sz, sz, vp, sz, sz, sz, vp, sz, vp, sz]),
    "rlg_solve_l_triangular_f32": (i32, [sz, sz, vp, sz, sz, sz, vp, sz, vp, sz]),
    "rlg_solve_triangular_f64_dev": (i32, [i32, sz, vp, sz, sz, vp, sz, vp, sz, vp]),
    "rlg_lu_decompose_f64": (i32, [sz, sz, vp, sz, C.POINTER(vp)]),
    "rlg_lu_decompose_f32": (i32, [sz, sz, vp, sz, C.POINTER(vp)]),
    "rlg_lu_decompose_f64_dev": (i32, [sz, vp, sz, vp, C.POINTER(vp)]),
    "rlg_lu_solve_f64": (i32, [vp, sz, sz, vp, sz, vp, sz]),
    "rlg_lu_solve_f32": (i32, [vp, sz, sz, vp, sz, vp, sz]),
    "rlg_lu_solve_f64_dev": (i32, [vp, sz, vp, sz, vp, sz, vp]),
    "rlg_lu_size": (i32, [vp, PSZ]),
    "rlg_lu_perm": (i32, [vp, vp]),
    "rlg_lu_ipiv": (i32, [vp, vp]),
    "rlg_lu_packed_f64": (i32, [vp, vp, sz]),
    "rlg_lu_packed_f32": (i32, [vp, vp, sz]),
    "rlg_lu_unpack_f64": (i32, [vp, vp, vp, vp]),
    "rlg_lu_det_f64": (i32, [vp, PD]),
    "rlg_lu_inverse_f64": (i32, [vp, vp, sz]),
    "rlg_permute_rows_f64": (i32, [sz, vp, sz, sz, vp, sz, vp, sz]),
    "rlg_permute_rows_f32": (i32, [sz, vp, sz, sz, vp, sz, vp, sz]),
    "rlg_permute_cols_f64": (i32, [sz, vp, sz, sz, vp, sz, vp, sz]),
    "rlg_permute_cols_f32": (i32, [sz, vp, sz, sz, vp, sz, vp, sz]),
    "rlg_permute_rows_f64_dev": (i32, [sz, vp, sz, vp, sz, vp, sz, vp]),
    "rlg_permute_cols_f64_dev": (i32, [sz, sz, vp, vp, sz, vp, sz, vp]),
    "rlg_lu_free": (None, [vp]),
    "rlg_lu_factor_f64_noalloc": (i32, [sz, vp, sz, vp, vp, vp]),
    "rlg_fplu_decompose_f64": (i32, [sz, sz, vp, sz, C.POINTER(vp)]),
    "rlg_fplu_decompose_f32": (i32, [sz, sz, vp, sz, C.POINTER(vp)]),
    "rlg_fplu_decompose_f64_dev": (i32, [sz, vp, sz, vp, C.POINTER(vp)]),
    "rlg_fplu_solve_f64": (i32, [vp, sz, sz, vp, sz, vp, sz]),
    "rlg_fplu_solve_f32": (i32, [vp, sz, sz, vp, sz, vp, sz]),
    "rlg_fplu_solve_true_f64": (i32, [vp, sz, sz, vp, sz, vp, sz]),
    "rlg_fplu_solve_f64_dev": (i32, [vp, sz, vp, sz, vp, sz, i32, vp]),
    "rlg_fplu_size": (i32, [vp, PSZ]),
    "rlg_fplu_perms": (i32, [vp, vp, vp]),
    "rlg_fplu_steps": (i32, [vp, PSZ]),
    "rlg_fplu_unpack_f64": (i32, [vp, vp, vp]),
    "rlg_fplu_unpack_f32": (i32, [vp, vp, vp]),
    "rlg_fplu_rank": (i32, [vp, PSZ]),
    "rlg_fplu_is_invertible": (i32, [vp, C.POINTER(i32)]),
    "rlg_fplu_det_f64": (i32, [vp, PD]),
    "rlg_fplu_inverse_f64": (i32, [vp, vp, sz]),
    "rlg_fplu_free": (None, [vp]),
    "rlg_chol_decompose_f64": (i32, [sz, sz, vp, sz, C.POINTER(vp)]),
    "rlg_chol_decompose_f32": (i32, [sz, sz, vp, sz, C.POINTER(vp)]),
    "rlg_chol_decompose_f64_dev": (i32, [sz, vp, sz, vp, C.POINTER(vp)]),
    "rlg_chol_solve_f64": (i32, [vp, sz, sz, vp, sz, vp, sz]),
    "rlg_chol_solve_f32": (i32, [vp, sz, sz, vp, sz, vp, sz]),
    "rlg_chol_solve_f64_dev": (i32, [vp, sz, vp, sz, vp, sz, vp]),
    "rlg_chol_size": (i32, [vp, PSZ]),
    "rlg_chol_unpack_f64": (i32, [vp, vp, sz]),
    "rlg_chol_unpack_f32": (i32, [vp, vp, sz]),
    "rlg_chol_det_f64": (i32, [vp, PD]),
    "rlg_chol_inverse_f64": (i32, [vp, vp, sz]),
    "rlg_chol_free": (None, [vp]),
    "rlg_mixed_opts_default": (None, [C.POINTER(MixedOpts)]),
    "rlg_solve_mixed_f64_dev": (i32, [sz, vp, sz, sz, vp, sz, vp, sz, C.POINTER(MixedOpts),
                                      C.POINTER(MixedReport), vp]),
    "rlg_solve_mixed_f64": (i32, [sz, vp, sz, sz, vp, sz, vp, sz, C.POINTER(MixedOpts),
                                  C.POINTER(MixedReport)]),
    "rlg_backward_error_f64_dev": (i32, [sz, vp, sz, sz, vp, sz, vp, sz, PD, vp]),
    "rlg_batch_solve_f64_dev": (i32, [sz, vp, vp]),
    "rlg_batch_solve_mixed_dev": (i32, [sz, vp, i32, dbl, PSZ, vp]),
    "rlg_batch_autotune": (i32, [i32, C.c_char_p, sz]),
    "rlg_batch_graph_create": (i32, [sz, vp, C.POINTER(vp)]),
    "rlg_batch_graph_launch": (i32, [vp, vp]),
    "rlg_batch_graph_destroy": (None, [vp]),
    "rlg_server_opts_default": (None, [C.POINTER(ServerOpts)]),
    "rlg_server_create": (i32, [C.POINTER(ServerOpts), C.POINTER(vp)]),
    "rlg_server_submit": (i32, [vp, C.POINTER(Request), PI64]),
    "rlg_server_wait": (i32, [vp, i64, C.POINTER(Result)]),
    "rlg_server_release": (i32, [vp, i64]),
    "rlg_server_evict": (i32, [vp, i64]),
    "rlg_server_destroy": (None, [vp]),
    "rlg_prof_enable": (i32, [i32]),
    "rlg_prof_reset": (i32, []),
    "rlg_prof_get": (i32, [C.c_char_p, PI64, PD, PD]),
    "rlg_prof_dump": (i32, [C.c_char_p, sz]),
}

_lib = None


def load():
    """Load librlgpu.so and attach the signatures of every exported entry point."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"librlgpu.so is not built ({LIB_PATH}); run `make -C {_HERE}` "
                "or __graft_entry__.build().  There is no CPU fallback.")
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(lib, name, None)
            if f is None:  # reported by the export test; calling it raises AttributeError
                continue
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


def last_error():
    msg = load().rlg_last_error_message()
    return msg.decode() if msg else ""


def ptr(a):
    """Address of a numpy array's data (or an int device pointer)."""
    if a is None:
        return None
    if isinstance(a, int):
        return a
    return a.ctypes.data
