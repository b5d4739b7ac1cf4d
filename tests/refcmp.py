"""The reference's elementwise comparators, restated for the tests.

assert_matrix_eq!/assert_vector_eq!/assert_scalar_eq! in rulinalg support
`exact`, `abs, tol`, `ulp, tol` and `float` (src/macros/comparison.rs:143-197,
src/ulp.rs:41-65).  `float` accepts |x - y| <= T::epsilon() or an ULP distance
of at most 4, where the ULP distance is the difference of the integer bit
patterns and is undefined (a failure) for NaN or differing signs.
"""
import numpy as np

EPS = {np.dtype(np.float64): np.finfo(np.float64).eps, np.dtype(np.float32): np.finfo(np.float32).eps}


def ulp_diff(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    it = np.int64 if a.dtype == np.float64 else np.int32
    ai = a.view(it).astype(np.int64)
    bi = b.view(it).astype(np.int64)
    d = np.abs(bi - ai).astype(np.float64)
    same = a == b
    bad = np.isnan(a) | np.isnan(b) | (np.signbit(a) != np.signbit(b))
    d = np.where(bad, np.inf, d)
    return np.where(same, 0.0, d)


def compare(x, y, comp="exact", tol=None, dtype=np.float64):
    x = np.asarray(x, dtype=dtype)
    y = np.asarray(y, dtype=dtype)
    assert x.shape == y.shape, (x.shape, y.shape)
    if comp == "exact":
        ok = x == y
    elif comp == "abs":
        ok = np.abs(x - y) <= tol
    elif comp == "ulp":
        ok = ulp_diff(x, y) <= tol
    elif comp == "float":
        ok = (np.abs(x - y) <= EPS[np.dtype(dtype)]) | (ulp_diff(x, y) <= 4)
    else:
        raise ValueError(comp)
    return bool(np.all(ok))


def assert_eq(x, y, comp="exact", tol=None, dtype=np.float64, what=""):
    assert compare(x, y, comp, tol, dtype), f"{what}: comp={comp} tol={tol}\n got {np.asarray(x)!r}\n exp {np.asarray(y)!r}"
