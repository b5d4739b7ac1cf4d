"""rulinalg_amd: the GPU-resident backend of rulinalg's GEMM and factor-and-solve
path on MI355X, seen from Python with the reference's names.

Every numeric call goes through librlgpu.so (include/rlgpu.h); this module only
moves numpy buffers across the C ABI and turns status codes into the
reference's error behaviour:

  reference                                        here
  ----------------------------------------------   ------------------------------------
  &a * &b  (src/matrix/mat_mul.rs:137-143)          a @ b / a * b on Matrix/MatrixSlice
  &m * &v  (src/matrix/impl_ops.rs:298-315)         m * v with a Vector
  PartialPivLu::decompose / solve / unpack / det   PartialPivLu.decompose(m).solve(b) ...
  / inverse (src/matrix/decomposition/lu.rs)
  FullPivLu::decompose / solve / unpack / det /    FullPivLu.decompose(m).solve(b) ...
  inverse / rank / is_invertible (lu.rs:339-599)
  Cholesky::decompose / solve / unpack / det /     Cholesky.decompose(m).solve(b) ...
  inverse (src/matrix/decomposition/cholesky.rs)
  Matrix::solve / inverse / det                    Matrix.solve / inverse / det
  (src/matrix/impl_mat.rs:354-427)
  Result::Err(Error{kind})  (src/error.rs:11-50)    raises Error with .kind() an ErrorKind
  assert!/panic on shapes                           raises ShapePanic (an AssertionError)
"""
import ctypes as C
import enum
from collections import namedtuple

import numpy as np

from . import _lib

__all__ = ["Matrix", "MatrixSlice", "Vector", "PermutationMatrix", "PartialPivLu", "LUP",
           "FullPivLu", "LUPQ",
           "Cholesky", "Error", "ErrorKind", "ShapePanic", "solve_mixed", "MixedOptions"]


class ErrorKind(enum.IntEnum):
    """ErrorKind, in the reference's declaration order (src/error.rs:21-34)."""
    InvalidArg = 1
    DecompFailure = 2
    AlgebraFailure = 3
    DivByZero = 4
    ScalarConversionFailure = 5
    InvalidPermutation = 6


class Error(Exception):
    """The reference's Error{kind, error} (src/error.rs:11-50)."""

    def __init__(self, kind, message=""):
        super().__init__(f"{kind.name}: {message}")
        self._kind = kind
        self.message = message

    def kind(self):
        return self._kind


class ShapePanic(AssertionError):
    """A reference `assert!` (dimension mismatch) that would panic."""


class DeviceError(RuntimeError):
    """A HIP runtime or kernel failure (no counterpart in the reference)."""


def _check(status):
    if status == 0:
        return
    msg = _lib.last_error()
    if 1 <= status <= 6:
        raise Error(ErrorKind(status), msg)
    if status == 7:
        raise ShapePanic(msg)
    raise DeviceError(f"{_lib.STATUS_NAMES.get(status, status)}: {msg}")


def _dtype_of(x):
    return np.float32 if np.asarray(x).dtype == np.float32 else np.float64


class Vector:
    """Vector<T> (src/vector/mod.rs:13-16): a length and a contiguous buffer."""

    def __init__(self, data, dtype=None):
        dt = dtype or _dtype_of(data)
        self._data = np.ascontiguousarray(np.asarray(data, dtype=dt).ravel())

    def size(self):
        return self._data.shape[0]

    def data(self):
        return self._data

    def into_vec(self):
        return list(self._data)

    def __getitem__(self, i):
        return self._data[i]

    def __len__(self):
        return self.size()

    def __repr__(self):
        return f"Vector({self._data!r})"


class _Operand:
    """Common part of Matrix and MatrixSlice: a strided row-major view."""

    def rows(self):
        return self._rows

    def cols(self):
        return self._cols

    def row_stride(self):
        return self._ld

    def _gemm(self, other):
        if self.dtype != other.dtype:
            raise TypeError("operands of different element types")
        out = np.empty((self._rows, other._cols), dtype=self.dtype)
        fn = _lib.load().rlg_gemm_f32 if self.dtype == np.float32 else _lib.load().rlg_gemm_f64
        _check(fn(self._rows, self._cols, self._base_ptr(), self._ld, other._rows, other._cols,
                  other._base_ptr(), other._ld, out.ctypes.data, max(other._cols, 1)))
        return Matrix(out)

    def __matmul__(self, other):
        return self.__mul__(other)

    def __mul__(self, other):
        if isinstance(other, _Operand):
            return self._gemm(other)
        if isinstance(other, PermutationMatrix):
            # M * P permutes columns: column j goes to perm[j] (impl_permutation_mul.rs:129-204)
            return other._permute(self, False)
        if isinstance(other, Vector):
            m = self.to_numpy()
            y = np.empty(self._rows)
            _check(_lib.load().rlg_matvec_f64(self._rows, self._cols, np.ascontiguousarray(m, float).ctypes.data,
                                              max(self._cols, 1), other.size(),
                                              np.ascontiguousarray(other.data(), float).ctypes.data,
                                              y.ctypes.data))
            return Vector(y)
        raise TypeError(f"unsupported operand {type(other)}")

    def _tri_solve(self, y, upper):
        if not isinstance(y, Vector):
            y = Vector(y)
        dt = self.dtype
        m = np.ascontiguousarray(self.to_numpy(), dtype=dt)
        b = np.ascontiguousarray(y.data(), dtype=dt)
        x = np.empty_like(b)
        lib = _lib.load()
        f32 = dt == np.float32
        if upper:
            fn = lib.rlg_solve_u_triangular_f32 if f32 else lib.rlg_solve_u_triangular_f64
        else:
            fn = lib.rlg_solve_l_triangular_f32 if f32 else lib.rlg_solve_l_triangular_f64
        n = max(self._cols, 1)
        _check(fn(self._rows, self._cols, m.ctypes.data, n, b.shape[0], 1, b.ctypes.data,
                  max(b.shape[0], 1), x.ctypes.data, max(b.shape[0], 1)))
        return Vector(x)

    def solve_u_triangular(self, y):
        """BaseMatrix::solve_u_triangular (src/matrix/base/mod.rs:1015-1024)."""
        return self._tri_solve(y, True)

    def solve_l_triangular(self, y):
        """BaseMatrix::solve_l_triangular (src/matrix/base/mod.rs:1058-1067)."""
        return self._tri_solve(y, False)


class Matrix(_Operand):
    """Matrix<T> (src/matrix/mod.rs:45-49): row-major, row_stride == cols."""

    def __init__(self, rows, cols=None, data=None, dtype=None):
        if cols is None:  # Matrix(ndarray)
            arr = np.asarray(rows)
            dt = dtype or _dtype_of(arr)
            arr = np.ascontiguousarray(arr, dtype=dt)
            if arr.ndim != 2:
                raise ValueError("Matrix needs a 2-D array")
        else:  # Matrix::new(rows, cols, data) (src/matrix/impl_mat.rs:31-41)
            dt = dtype or _dtype_of(data)
            flat = np.asarray(data, dtype=dt).ravel()
            if flat.shape[0] != rows * cols:
                raise ShapePanic("Data does not match given dimensions.")
            arr = np.ascontiguousarray(flat.reshape(rows, cols))
        self._a = arr
        self._rows, self._cols = arr.shape
        self._ld = max(self._cols, 1)
        self.dtype = arr.dtype

    @staticmethod
    def identity(n, dtype=np.float64):
        return Matrix(np.eye(n, dtype=dtype))

    @staticmethod
    def zeros(rows, cols, dtype=np.float64):
        return Matrix(np.zeros((rows, cols), dtype=dtype))

    def _base_ptr(self):
        return self._a.ctypes.data

    def data(self):
        return self._a.ravel()

    def to_numpy(self):
        return self._a

    def into_vec(self):
        return list(self._a.ravel())

    def __getitem__(self, ij):
        i, j = ij
        return self._a[i, j]

    def __eq__(self, other):
        return isinstance(other, Matrix) and np.array_equal(self._a, other._a)

    def __repr__(self):
        return f"Matrix({self._a!r})"

    # Matrix::solve / inverse / det (src/matrix/impl_mat.rs:354-427)
    def solve(self, y):
        return PartialPivLu.decompose(self).solve(y)

    def inverse(self):
        return PartialPivLu.decompose(self).inverse()

    def det(self):
        """Matrix::det (impl_mat.rs:407-427): closed forms for diagonal, 2x2 and
        3x3 matrices, otherwise PartialPivLu (zero when it fails)."""
        if self._rows != self._cols:
            raise ShapePanic("Matrix is not square.")
        a = self._a
        n = self._cols
        if np.count_nonzero(a - np.diag(np.diag(a))) == 0:
            d = a.dtype.type(1)
            for v in np.diag(a):
                d = d * v
            return d
        if n == 2:
            return (a[0, 0] * a[1, 1]) - (a[0, 1] * a[1, 0])
        if n == 3:
            return ((a[0, 0] * a[1, 1] * a[2, 2]) + (a[0, 1] * a[1, 2] * a[2, 0]) +
                    (a[0, 2] * a[1, 0] * a[2, 1]) - (a[0, 0] * a[1, 2] * a[2, 1]) -
                    (a[0, 1] * a[1, 0] * a[2, 2]) - (a[0, 2] * a[1, 1] * a[2, 0]))
        try:
            return PartialPivLu.decompose(self).det()
        except Error:
            return a.dtype.type(0)


class MatrixSlice(_Operand):
    """MatrixSlice::from_matrix (src/matrix/slice.rs:15-33): a view with
    row_stride = parent.cols, usable as a GEMM operand without copying."""

    def __init__(self, mat, start, rows, cols):
        r0, c0 = start
        if r0 + rows > mat.rows() or c0 + cols > mat.cols():
            raise ShapePanic("View dimensions exceed matrix dimensions.")
        self._parent = mat
        self._off = r0 * mat.row_stride() + c0
        self._rows, self._cols = rows, cols
        self._ld = mat.row_stride()
        self.dtype = mat.dtype

    @staticmethod
    def from_matrix(mat, start, rows, cols):
        return MatrixSlice(mat, start, rows, cols)

    def _base_ptr(self):
        return self._parent._a.ctypes.data + self._off * self.dtype.itemsize

    def to_numpy(self):
        r0, c0 = divmod(self._off, self._ld)
        return self._parent._a[r0:r0 + self._rows, c0:c0 + self._cols]


class PermutationMatrix:
    """PermutationMatrix<T> (src/matrix/permutation_matrix.rs:103-111)."""

    def __init__(self, perm):
        self._perm = np.asarray(perm, dtype=np.int64)

    def size(self):
        return len(self._perm)

    def into_vec(self):
        return list(int(p) for p in self._perm)

    def inverse(self):
        inv = np.empty_like(self._perm)
        inv[self._perm] = np.arange(len(self._perm))
        return PermutationMatrix(inv)

    def as_matrix(self):
        # Matrix::from_fn passes (col, row) (impl_mat.rs:74-84): column c has its 1 at row perm[c]
        n = len(self._perm)
        m = np.zeros((n, n))
        m[self._perm, np.arange(n)] = 1.0
        return Matrix(m)

    def _permute(self, m, by_rows):
        """Row (P * M) or column (M * P) permutation of a Matrix on the GPU
        (rlg_permute_rows / rlg_permute_cols)."""
        a = np.ascontiguousarray(m.to_numpy())
        out = np.empty_like(a)
        perm = np.ascontiguousarray(self._perm, dtype=np.uint64)
        lib = _lib.load()
        f32 = a.dtype == np.float32
        if by_rows:
            fn = lib.rlg_permute_rows_f32 if f32 else lib.rlg_permute_rows_f64
        else:
            fn = lib.rlg_permute_cols_f32 if f32 else lib.rlg_permute_cols_f64
        ld = max(a.shape[1], 1)
        _check(fn(len(perm), perm.ctypes.data, a.shape[0], a.shape[1], a.ctypes.data, ld,
                  out.ctypes.data, ld))
        return Matrix(out)

    def __mul__(self, other):
        # P * Matrix permutes rows: row i goes to perm[i] (impl_permutation_mul.rs:53-127)
        if isinstance(other, (Matrix, MatrixSlice)):
            return self._permute(other, True)
        if isinstance(other, Vector):
            out = np.empty_like(other.data())
            out[self._perm] = other.data()
            return Vector(out)
        raise TypeError(type(other))

    def __eq__(self, other):
        return isinstance(other, PermutationMatrix) and np.array_equal(self._perm, other._perm)


LUP = namedtuple("LUP", ["l", "u", "p"])


class PartialPivLu:
    """PartialPivLu<T> (src/matrix/decomposition/lu.rs:130-300), factor on the GPU."""

    def __init__(self, handle, n, dtype):
        self._h = handle
        self._n = n
        self.dtype = dtype

    @staticmethod
    def decompose(matrix):
        m = matrix if isinstance(matrix, Matrix) else Matrix(matrix)
        h = C.c_void_p()
        lib = _lib.load()
        fn = lib.rlg_lu_decompose_f32 if m.dtype == np.float32 else lib.rlg_lu_decompose_f64
        _check(fn(m.rows(), m.cols(), m._base_ptr(), m.row_stride(), C.byref(h)))
        return PartialPivLu(h, m.rows(), m.dtype)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            _lib.load().rlg_lu_free(h)
            self._h = None

    def solve(self, b):
        bv = b if isinstance(b, Vector) else Vector(b, dtype=self.dtype)
        data = np.ascontiguousarray(bv.data(), dtype=self.dtype)
        x = np.empty_like(data)
        lib = _lib.load()
        fn = lib.rlg_lu_solve_f32 if self.dtype == np.float32 else lib.rlg_lu_solve_f64
        _check(fn(self._h, data.shape[0], 1, data.ctypes.data, max(data.shape[0], 1), x.ctypes.data,
                  max(data.shape[0], 1)))
        return Vector(x)

    def solve_many(self, b):
        """nrhs right-hand sides at once: b is (nrhs, n), one vector per row."""
        bm = np.ascontiguousarray(b, dtype=self.dtype)
        x = np.empty_like(bm)
        lib = _lib.load()
        fn = lib.rlg_lu_solve_f32 if self.dtype == np.float32 else lib.rlg_lu_solve_f64
        _check(fn(self._h, bm.shape[1], bm.shape[0], bm.ctypes.data, bm.shape[1], x.ctypes.data,
                  bm.shape[1]))
        return x

    def packed(self):
        out = np.empty((self._n, self._n), dtype=self.dtype)
        lib = _lib.load()
        fn = lib.rlg_lu_packed_f32 if self.dtype == np.float32 else lib.rlg_lu_packed_f64
        _check(fn(self._h, out.ctypes.data, max(self._n, 1)))
        return out

    def perm(self):
        p = np.empty(self._n, dtype=np.uintp)
        _check(_lib.load().rlg_lu_perm(self._h, p.ctypes.data))
        return p.astype(np.int64)

    def ipiv(self):
        p = np.empty(self._n, dtype=np.int64)
        _check(_lib.load().rlg_lu_ipiv(self._h, p.ctypes.data))
        return p

    def unpack(self):
        """Decomposition::unpack (lu.rs:135-150)."""
        lu = self.packed()
        n = self._n
        l = np.tril(lu, -1) + np.eye(n, dtype=lu.dtype)
        u = np.triu(lu)
        return LUP(Matrix(l), Matrix(u), PermutationMatrix(self.perm()))

    def det(self):
        d = C.c_double()
        _check(_lib.load().rlg_lu_det_f64(self._h, C.byref(d)))
        return d.value

    def inverse(self):
        inv = np.empty((self._n, self._n))
        _check(_lib.load().rlg_lu_inverse_f64(self._h, inv.ctypes.data, max(self._n, 1)))
        return Matrix(inv)


LUPQ = namedtuple("LUPQ", ["l", "u", "p", "q"])


class FullPivLu:
    """FullPivLu<T> (src/matrix/decomposition/lu.rs:339-599), factor on the GPU.

    solve() returns the reference's result (lu.rs:458-464), which maps the
    substitutions' result through `&q * y`: that is the solution of A x = b
    when q is its own inverse and a permutation of it otherwise.
    solve_true() maps through q^-1 and returns the solution for every q."""

    def __init__(self, handle, n, dtype):
        self._h = handle
        self._n = n
        self.dtype = dtype

    @staticmethod
    def decompose(matrix):
        m = matrix if isinstance(matrix, Matrix) else Matrix(matrix)
        h = C.c_void_p()
        lib = _lib.load()
        fn = lib.rlg_fplu_decompose_f32 if m.dtype == np.float32 else lib.rlg_fplu_decompose_f64
        _check(fn(m.rows(), m.cols(), m._base_ptr(), m.row_stride(), C.byref(h)))
        return FullPivLu(h, m.rows(), m.dtype)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            _lib.load().rlg_fplu_free(h)
            self._h = None

    def _solve(self, b, true_solution):
        bm = np.ascontiguousarray(b, dtype=self.dtype)
        one = bm.ndim == 1
        bm = bm.reshape(1, -1) if one else bm
        x = np.empty_like(bm)
        lib = _lib.load()
        if true_solution:
            fn = lib.rlg_fplu_solve_true_f64
        else:
            fn = lib.rlg_fplu_solve_f32 if self.dtype == np.float32 else lib.rlg_fplu_solve_f64
        ld = max(bm.shape[1], 1)
        _check(fn(self._h, bm.shape[1], bm.shape[0], bm.ctypes.data, ld, x.ctypes.data, ld))
        return x[0] if one else x

    def solve(self, b):
        bv = b if isinstance(b, Vector) else Vector(b, dtype=self.dtype)
        return Vector(self._solve(bv.data(), False))

    def solve_many(self, b):
        """nrhs right-hand sides at once: b is (nrhs, n), one vector per row."""
        return self._solve(b, False)

    def solve_true(self, b):
        bv = b if isinstance(b, Vector) else Vector(b)
        return Vector(self._solve(bv.data(), True))

    def perms(self):
        p = np.empty(self._n, dtype=np.uintp)
        q = np.empty(self._n, dtype=np.uintp)
        _check(_lib.load().rlg_fplu_perms(self._h, p.ctypes.data, q.ctypes.data))
        return p.astype(np.int64), q.astype(np.int64)

    def steps(self):
        s = C.c_size_t()
        _check(_lib.load().rlg_fplu_steps(self._h, C.byref(s)))
        return s.value

    def unpack(self):
        """Decomposition::unpack (lu.rs:348-361)."""
        n = self._n
        l = np.empty((n, n), dtype=self.dtype)
        u = np.empty((n, n), dtype=self.dtype)
        lib = _lib.load()
        fn = lib.rlg_fplu_unpack_f32 if self.dtype == np.float32 else lib.rlg_fplu_unpack_f64
        _check(fn(self._h, l.ctypes.data, u.ctypes.data))
        p, q = self.perms()
        return LUPQ(Matrix(l), Matrix(u), PermutationMatrix(p), PermutationMatrix(q))

    def rank(self):
        r = C.c_size_t()
        _check(_lib.load().rlg_fplu_rank(self._h, C.byref(r)))
        return r.value

    def is_invertible(self):
        y = C.c_int()
        _check(_lib.load().rlg_fplu_is_invertible(self._h, C.byref(y)))
        return bool(y.value)

    def det(self):
        d = C.c_double()
        _check(_lib.load().rlg_fplu_det_f64(self._h, C.byref(d)))
        return d.value

    def inverse(self):
        inv = np.empty((self._n, self._n))
        _check(_lib.load().rlg_fplu_inverse_f64(self._h, inv.ctypes.data, max(self._n, 1)))
        return Matrix(inv)


class Cholesky:
    """Cholesky<T> (src/matrix/decomposition/cholesky.rs:94-246), factor on the GPU."""

    def __init__(self, handle, n, dtype):
        self._h = handle
        self._n = n
        self.dtype = dtype

    @staticmethod
    def decompose(matrix):
        m = matrix if isinstance(matrix, Matrix) else Matrix(matrix)
        h = C.c_void_p()
        lib = _lib.load()
        fn = lib.rlg_chol_decompose_f32 if m.dtype == np.float32 else lib.rlg_chol_decompose_f64
        _check(fn(m.rows(), m.cols(), m._base_ptr(), m.row_stride(), C.byref(h)))
        return Cholesky(h, m.rows(), m.dtype)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            _lib.load().rlg_chol_free(h)
            self._h = None

    def solve(self, b):
        bv = b if isinstance(b, Vector) else Vector(b, dtype=self.dtype)
        data = np.ascontiguousarray(bv.data(), dtype=self.dtype)
        x = np.empty_like(data)
        lib = _lib.load()
        fn = lib.rlg_chol_solve_f32 if self.dtype == np.float32 else lib.rlg_chol_solve_f64
        _check(fn(self._h, data.shape[0], 1, data.ctypes.data, max(data.shape[0], 1), x.ctypes.data,
                  max(data.shape[0], 1)))
        return Vector(x)

    def solve_many(self, b):
        bm = np.ascontiguousarray(b, dtype=self.dtype)
        x = np.empty_like(bm)
        lib = _lib.load()
        fn = lib.rlg_chol_solve_f32 if self.dtype == np.float32 else lib.rlg_chol_solve_f64
        _check(fn(self._h, bm.shape[1], bm.shape[0], bm.ctypes.data, bm.shape[1], x.ctypes.data,
                  bm.shape[1]))
        return x

    def unpack(self):
        """Decomposition::unpack (cholesky.rs:237-246): L with zeros above."""
        out = np.empty((self._n, self._n), dtype=self.dtype)
        lib = _lib.load()
        fn = lib.rlg_chol_unpack_f32 if self.dtype == np.float32 else lib.rlg_chol_unpack_f64
        _check(fn(self._h, out.ctypes.data, max(self._n, 1)))
        return Matrix(out)

    def det(self):
        d = C.c_double()
        _check(_lib.load().rlg_chol_det_f64(self._h, C.byref(d)))
        return d.value

    def inverse(self):
        inv = np.empty((self._n, self._n))
        _check(_lib.load().rlg_chol_inverse_f64(self._h, inv.ctypes.data, max(self._n, 1)))
        return Matrix(inv)


class MixedOptions:
    """Options of the mixed-precision refined solve (include/rlgpu.h)."""

    def __init__(self, factor_prec=_lib.PREC_F16X3, refine=_lib.REFINE_IR, tol=1e-13, max_iter=30,
                 fallback=True, pivot=_lib.PIVOT_AUTO):
        self.factor_prec, self.refine, self.tol = factor_prec, refine, tol
        self.max_iter, self.fallback, self.pivot = max_iter, fallback, pivot

    def to_c(self):
        return _lib.MixedOpts(self.factor_prec, self.refine, self.tol, self.max_iter,
                              1 if self.fallback else 0, self.pivot)


def solve_mixed(a, b, options=None):
    """Solve A X = B (B: (nrhs, n) or (n,)) by a low-precision factorization and
    FP64 refinement.  Returns (X, report dict)."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    bb = np.ascontiguousarray(np.atleast_2d(b), dtype=np.float64)
    n = a.shape[0]
    if a.shape != (n, n) or bb.shape[1] != n:
        raise ShapePanic("dimension mismatch")
    x = np.empty_like(bb)
    opts = (options or MixedOptions()).to_c()
    rep = _lib.MixedReport()
    _check(_lib.load().rlg_solve_mixed_f64(n, a.ctypes.data, n, bb.shape[0], bb.ctypes.data, n,
                                           x.ctypes.data, n, C.byref(opts), C.byref(rep)))
    report = {f: getattr(rep, f) for f, _ in _lib.MixedReport._fields_}
    return (x[0] if np.ndim(b) == 1 else x), report
