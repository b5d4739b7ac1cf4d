"""ctypes front end of the C oracle in rlg_oracle.c (test infrastructure only).

Each wrapper names the reference routine it restates; see the C source for the
file:line citations.  Arrays are numpy, row-major, float64 unless stated.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

OK, E_INVALID_ARG, E_DECOMP_FAILURE, E_ALGEBRA_FAILURE, E_DIV_BY_ZERO = 0, 1, 2, 3, 4
E_SHAPE = 7

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: run `make -C {_HERE}`")
        _lib = C.CDLL(_LIB_PATH)
        sz, i64, i32 = C.c_size_t, C.c_int64, C.c_int
        for suf, ct in (("f64", C.c_double), ("f32", C.c_float)):
            P = C.POINTER(ct)
            f = getattr(_lib, f"orc_dot_{suf}"); f.restype = ct; f.argtypes = [P, P, sz]
            f = getattr(_lib, f"orc_matmul_{suf}"); f.restype = i32
            f.argtypes = [sz, sz, sz, sz, P, sz, P, sz, P]
            f = getattr(_lib, f"orc_matvec_{suf}"); f.restype = None; f.argtypes = [sz, sz, P, P, P]
            f = getattr(_lib, f"orc_lu_decompose_{suf}"); f.restype = i32
            f.argtypes = [sz, P, C.POINTER(sz), C.POINTER(i64), P, C.POINTER(sz)]
            f = getattr(_lib, f"orc_lu_solve_{suf}"); f.restype = i32
            f.argtypes = [sz, P, C.POINTER(sz), P, P]
            f = getattr(_lib, f"orc_lu_forward_substitution_{suf}"); f.restype = None
            f.argtypes = [sz, P, P]
            for nm in ("back_substitution", "forward_substitution", "transpose_back_substitution"):
                f = getattr(_lib, f"orc_{nm}_{suf}"); f.restype = i32; f.argtypes = [sz, P, sz, P]
            f = getattr(_lib, f"orc_lu_det_{suf}"); f.restype = ct; f.argtypes = [sz, P, C.POINTER(sz)]
            f = getattr(_lib, f"orc_lu_inverse_{suf}"); f.restype = i32
            f.argtypes = [sz, P, C.POINTER(sz), P]
            f = getattr(_lib, f"orc_chol_decompose_{suf}"); f.restype = i32
            f.argtypes = [sz, P, C.POINTER(sz)]
            f = getattr(_lib, f"orc_chol_solve_{suf}"); f.restype = i32; f.argtypes = [sz, P, P, P]
            f = getattr(_lib, f"orc_chol_det_{suf}"); f.restype = ct; f.argtypes = [sz, P]
            f = getattr(_lib, f"orc_chol_inverse_{suf}"); f.restype = i32; f.argtypes = [sz, P, P]
            f = getattr(_lib, f"orc_lu_unpack_{suf}"); f.restype = None; f.argtypes = [sz, P, P, P]
            f = getattr(_lib, f"orc_chol_unpack_{suf}"); f.restype = None; f.argtypes = [sz, P]
            f = getattr(_lib, f"orc_fplu_decompose_{suf}"); f.restype = i32
            f.argtypes = [sz, P, C.POINTER(sz), C.POINTER(sz), C.POINTER(i64), C.POINTER(i64),
                          C.POINTER(sz)]
            f = getattr(_lib, f"orc_fplu_solve_{suf}"); f.restype = i32
            f.argtypes = [sz, P, C.POINTER(sz), C.POINTER(sz), P, P]
            f = getattr(_lib, f"orc_fplu_rank_{suf}"); f.restype = sz; f.argtypes = [sz, P]
            f = getattr(_lib, f"orc_fplu_is_invertible_{suf}"); f.restype = i32; f.argtypes = [sz, P]
            f = getattr(_lib, f"orc_fplu_det_{suf}"); f.restype = ct
            f.argtypes = [sz, P, C.POINTER(sz), C.POINTER(sz)]
            f = getattr(_lib, f"orc_fplu_inverse_{suf}"); f.restype = i32
            f.argtypes = [sz, P, C.POINTER(sz), C.POINTER(sz), P]
        _lib.orc_perm_inverse.restype = None
        _lib.orc_perm_inverse.argtypes = [sz, C.POINTER(sz), C.POINTER(sz)]
        _lib.orc_perm_parity.restype = i32
        _lib.orc_perm_parity.argtypes = [sz, C.POINTER(sz)]
        _lib.orc_gemm_blocked_f64.restype = None
        _lib.orc_gemm_blocked_f64.argtypes = [sz, sz, sz, C.POINTER(C.c_double), sz,
                                              C.POINTER(C.c_double), sz, C.POINTER(C.c_double)]
    return _lib


def _suf(dtype):
    return "f32" if np.dtype(dtype) == np.float32 else "f64"


def _ptr(a):
    ct = C.c_float if a.dtype == np.float32 else C.c_double
    return a.ctypes.data_as(C.POINTER(ct))


def _szp(a):
    return a.ctypes.data_as(C.POINTER(C.c_size_t))


def _f(a, dtype=np.float64):
    return np.ascontiguousarray(a, dtype=dtype)


class OracleError(Exception):
    def __init__(self, code, where=None):
        super().__init__(f"oracle status {code}" + ("" if where is None else f" at {where}"))
        self.code = code
        self.where = where


def dot(u, v):
    """utils::dot (src/utils.rs:20-51)."""
    dt = np.asarray(u).dtype if np.asarray(u).dtype == np.float32 else np.float64
    u, v = _f(u, dt), _f(v, dt)
    return getattr(lib(), f"orc_dot_{_suf(dt)}")(_ptr(u), _ptr(v), min(len(u), len(v)))


def matmul(a, b, lda=None, ldb=None, rows=None, inner=None, cols=None):
    """Generic product of mat_mul_general! (src/matrix/mat_mul.rs:76-98), i-k-j order.

    With lda/ldb the operands are strided views of the given buffers
    (MatrixSlice semantics, src/matrix/slice.rs:15-33)."""
    dt = np.float32 if np.asarray(a).dtype == np.float32 else np.float64
    a, b = _f(a, dt), _f(b, dt)
    if lda is None:
        p, q = a.shape
        lda = q
    else:
        p, q = rows, inner
    if ldb is None:
        q2, r = b.shape
        ldb = r
    else:
        q2, r = inner, cols
    c = np.empty((p, r), dtype=dt)
    st = getattr(lib(), f"orc_matmul_{_suf(dt)}")(p, q, q2, r, _ptr(a), lda, _ptr(b), ldb, _ptr(c))
    if st != OK:
        raise OracleError(st)
    return c


def matvec(a, v):
    """&Matrix * &Vector (src/matrix/impl_ops.rs:298-315)."""
    a, v = _f(a), _f(v)
    y = np.empty(a.shape[0])
    lib().orc_matvec_f64(a.shape[0], a.shape[1], _ptr(a), _ptr(v), _ptr(y))
    return y


def lu_decompose(a, dtype=np.float64, with_margin=False):
    """PartialPivLu::decompose (src/matrix/decomposition/lu.rs:163-194).

    Returns (packed lu, stored perm p, step pivots ipiv[, margins]).  Raises
    OracleError(E_DIV_BY_ZERO, step) on a pivot below epsilon."""
    lu = _f(a, dtype).copy()
    n = lu.shape[0]
    if lu.ndim != 2 or lu.shape[1] != n:
        raise OracleError(E_SHAPE)
    perm = np.zeros(n, dtype=np.uintp)
    ipiv = np.zeros(n, dtype=np.int64)
    margin = np.zeros(n, dtype=dtype) if with_margin else None
    fail = C.c_size_t(0)
    st = getattr(lib(), f"orc_lu_decompose_{_suf(dtype)}")(
        n, _ptr(lu), _szp(perm), ipiv.ctypes.data_as(C.POINTER(C.c_int64)),
        _ptr(margin) if with_margin else None, C.byref(fail))
    if st != OK:
        raise OracleError(st, fail.value)
    out = (lu, perm.astype(np.int64), ipiv)
    return out + (margin,) if with_margin else out


def lu_solve(lu, perm, b):
    """PartialPivLu::solve (src/matrix/decomposition/lu.rs:231-244)."""
    dt = lu.dtype
    lu, b = _f(lu, dt), _f(b, dt)
    p = np.ascontiguousarray(perm, dtype=np.uintp)
    x = np.empty_like(b)
    st = getattr(lib(), f"orc_lu_solve_{_suf(dt)}")(lu.shape[0], _ptr(lu), _szp(p), _ptr(b), _ptr(x))
    if st != OK:
        raise OracleError(st)
    return x


def lu_forward_substitution(lu, b):
    """lu_forward_substitution (src/matrix/decomposition/lu.rs:624-642)."""
    lu, x = _f(lu), _f(b).copy()
    lib().orc_lu_forward_substitution_f64(lu.shape[0], _ptr(lu), _ptr(x))
    return x


def _subst(name, m, b):
    m, x = _f(m), _f(b).copy()
    st = getattr(lib(), f"orc_{name}_f64")(m.shape[0], _ptr(m), m.shape[1] if m.size else 0, _ptr(x))
    if st != OK:
        raise OracleError(st)
    return x


def back_substitution(u, b):
    """back_substitution (src/matrix/mod.rs:318-357)."""
    return _subst("back_substitution", u, b)


def forward_substitution(l, b):
    """forward_substitution (src/matrix/mod.rs:363-398)."""
    return _subst("forward_substitution", l, b)


def transpose_back_substitution(l, b):
    """transpose_back_substitution (src/matrix/decomposition/cholesky.rs:329-361)."""
    return _subst("transpose_back_substitution", l, b)


def lu_det(lu, perm):
    """PartialPivLu::det (src/matrix/decomposition/lu.rs:291-300)."""
    lu = _f(lu, lu.dtype)
    p = np.ascontiguousarray(perm, dtype=np.uintp)
    return getattr(lib(), f"orc_lu_det_{_suf(lu.dtype)}")(lu.shape[0], _ptr(lu), _szp(p))


def lu_inverse(lu, perm):
    """PartialPivLu::inverse (src/matrix/decomposition/lu.rs:251-285)."""
    lu = _f(lu)
    p = np.ascontiguousarray(perm, dtype=np.uintp)
    inv = np.empty_like(lu)
    st = lib().orc_lu_inverse_f64(lu.shape[0], _ptr(lu), _szp(p), _ptr(inv))
    if st != OK:
        raise OracleError(st)
    return inv


def lu_unpack(lu):
    """PartialPivLu::unpack (src/matrix/decomposition/lu.rs:135-150) -> (L, U)."""
    lu = _f(lu)
    l, u = np.empty_like(lu), np.empty_like(lu)
    lib().orc_lu_unpack_f64(lu.shape[0], _ptr(lu), _ptr(l), _ptr(u))
    return l, u


def perm_matrix(perm):
    """PermutationMatrix::as_matrix (src/matrix/permutation_matrix.rs:239-247).

    Matrix::from_fn passes (col, row) to its closure (src/matrix/impl_mat.rs:74-84),
    so column c holds its one at row perm[c]; M*v then puts v[c] at perm[c]."""
    n = len(perm)
    m = np.zeros((n, n))
    m[np.asarray(perm, dtype=np.int64), np.arange(n)] = 1.0
    return m


def permute_rows(perm, a):
    """PermutationMatrix * Matrix through permute_rows_into_buffer
    (src/matrix/permutation_matrix.rs:318-331): source row i is copied to
    buffer row perm[i]."""
    a = np.asarray(a)
    out = np.zeros_like(a)
    for i, t in enumerate(perm):
        out[int(t), :] = a[i, :]
    return out


def permute_cols(perm, a):
    """Matrix * PermutationMatrix through permute_cols_into_buffer
    (src/matrix/permutation_matrix.rs:342-358): in every row, source element j
    is copied to position perm[j]."""
    a = np.asarray(a)
    out = np.zeros_like(a)
    for r in range(a.shape[0]):
        for j, t in enumerate(perm):
            out[r, int(t)] = a[r, j]
    return out


def perm_inverse(perm):
    """PermutationMatrix::inverse (src/matrix/permutation_matrix.rs:137-148)."""
    p = np.ascontiguousarray(perm, dtype=np.uintp)
    inv = np.empty_like(p)
    lib().orc_perm_inverse(len(p), _szp(p), _szp(inv))
    return inv.astype(np.int64)


def perm_parity(perm):
    """PermutationMatrix::parity (src/matrix/permutation_matrix.rs:216-234): +1 even, -1 odd."""
    p = np.ascontiguousarray(perm, dtype=np.uintp)
    return lib().orc_perm_parity(len(p), _szp(p))


def chol_decompose(a, dtype=np.float64):
    """Cholesky::decompose (src/matrix/decomposition/cholesky.rs:116-169).

    Returns the working matrix: L in the lower triangle, the input above."""
    w = _f(a, dtype).copy()
    n = w.shape[0]
    if w.ndim != 2 or w.shape[1] != n:
        raise OracleError(E_SHAPE)
    fail = C.c_size_t(0)
    st = getattr(lib(), f"orc_chol_decompose_{_suf(dtype)}")(n, _ptr(w), C.byref(fail))
    if st != OK:
        raise OracleError(st, fail.value)
    return w


def chol_unpack(w):
    """Cholesky::unpack (src/matrix/decomposition/cholesky.rs:237-246)."""
    l = _f(w, w.dtype).copy()
    getattr(lib(), f"orc_chol_unpack_{_suf(l.dtype)}")(l.shape[0], _ptr(l))
    return l


def chol_solve(w, b):
    """Cholesky::solve (src/matrix/decomposition/cholesky.rs:194-202)."""
    dt = w.dtype
    w, b = _f(w, dt), _f(b, dt)
    x = np.empty_like(b)
    st = getattr(lib(), f"orc_chol_solve_{_suf(dt)}")(w.shape[0], _ptr(w), _ptr(b), _ptr(x))
    if st != OK:
        raise OracleError(st)
    return x


def chol_det(w):
    """Cholesky::det (src/matrix/decomposition/cholesky.rs:175-180)."""
    w = _f(w)
    return lib().orc_chol_det_f64(w.shape[0], _ptr(w))


def chol_inverse(w):
    """Cholesky::inverse (src/matrix/decomposition/cholesky.rs:209-234)."""
    w = _f(w)
    inv = np.empty_like(w)
    st = lib().orc_chol_inverse_f64(w.shape[0], _ptr(w), _ptr(inv))
    if st != OK:
        raise OracleError(st)
    return inv


def gemm_blocked(a, b):
    """Packed, blocked single-thread DGEMM: the CPU timing stand-in for matrixmultiply::dgemm."""
    a, b = _f(a), _f(b)
    m, k = a.shape
    n = b.shape[1]
    c = np.empty((m, n))
    lib().orc_gemm_blocked_f64(m, k, n, _ptr(a), k, _ptr(b), n, _ptr(c))
    return c


def fplu_decompose(a, dtype=np.float64):
    """FullPivLu::decompose (src/matrix/decomposition/lu.rs:383-432).

    Returns (packed lu, stored p, stored q, step pivot rows, step pivot
    columns, steps taken); the step arrays hold the first `steps` entries."""
    lu = _f(a, dtype).copy()
    n = lu.shape[0] if lu.ndim == 2 else 0
    if lu.ndim != 2 or lu.shape[1] != n:
        raise OracleError(E_SHAPE)
    p = np.zeros(n, dtype=np.uintp)
    q = np.zeros(n, dtype=np.uintp)
    prow = np.zeros(max(n, 1), dtype=np.int64)
    pcol = np.zeros(max(n, 1), dtype=np.int64)
    steps = C.c_size_t(0)
    st = getattr(lib(), f"orc_fplu_decompose_{_suf(dtype)}")(
        n, _ptr(lu), _szp(p), _szp(q), prow.ctypes.data_as(C.POINTER(C.c_int64)),
        pcol.ctypes.data_as(C.POINTER(C.c_int64)), C.byref(steps))
    if st != OK:
        raise OracleError(st)
    k = steps.value
    return lu, p.astype(np.int64), q.astype(np.int64), prow[:k].copy(), pcol[:k].copy(), k


def _fplu_args(lu, p, q):
    dt = lu.dtype
    return (_f(lu, dt), np.ascontiguousarray(p, dtype=np.uintp), np.ascontiguousarray(q, dtype=np.uintp))


def fplu_solve(lu, p, q, b):
    """FullPivLu::solve (src/matrix/decomposition/lu.rs:458-464)."""
    lu, p, q = _fplu_args(lu, p, q)
    b = _f(b, lu.dtype)
    x = np.empty_like(b)
    st = getattr(lib(), f"orc_fplu_solve_{_suf(lu.dtype)}")(lu.shape[0], _ptr(lu), _szp(p), _szp(q),
                                                           _ptr(b), _ptr(x))
    if st != OK:
        raise OracleError(st)
    return x


def fplu_rank(lu):
    """FullPivLu::rank (src/matrix/decomposition/lu.rs:544-557)."""
    lu = _f(lu, lu.dtype)
    return getattr(lib(), f"orc_fplu_rank_{_suf(lu.dtype)}")(lu.shape[0], _ptr(lu))


def fplu_is_invertible(lu):
    """FullPivLu::is_invertible (src/matrix/decomposition/lu.rs:582-594)."""
    lu = _f(lu, lu.dtype)
    return bool(getattr(lib(), f"orc_fplu_is_invertible_{_suf(lu.dtype)}")(lu.shape[0], _ptr(lu)))


def fplu_det(lu, p, q):
    """FullPivLu::det (src/matrix/decomposition/lu.rs:505-517)."""
    lu, p, q = _fplu_args(lu, p, q)
    return getattr(lib(), f"orc_fplu_det_{_suf(lu.dtype)}")(lu.shape[0], _ptr(lu), _szp(p), _szp(q))


def fplu_inverse(lu, p, q):
    """FullPivLu::inverse (src/matrix/decomposition/lu.rs:472-497)."""
    lu, p, q = _fplu_args(lu, p, q)
    inv = np.empty_like(lu)
    st = getattr(lib(), f"orc_fplu_inverse_{_suf(lu.dtype)}")(lu.shape[0], _ptr(lu), _szp(p), _szp(q),
                                                             _ptr(inv))
    if st != OK:
        raise OracleError(st)
    return inv
