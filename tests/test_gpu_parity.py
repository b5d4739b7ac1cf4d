"""GPU parity: librlgpu (through its C ABI) against the CPU oracle.

Tolerances (stated here, used below):
  * bit-exact: the reference's `exact` KATs; the packed LU factor, pivot
    sequence and Cholesky factor for orders <= 64 (one panel: the kernels use
    the reference's operation order); pivot sequences at every order where the
    oracle's pivot margin is clear of rounding (no near-ties).
  * GEMM: ||C - C_oracle||_F / (||A||_F ||B||_F) <= 1e-13 and the elementwise
    bound |C - C_exact| <= gamma_k |A||B|.
  * solves: normwise backward error <= 1e-13, and
    ||x - x_oracle||inf / ||x_oracle||inf <= max(1e-13, 16 n kappa_inf u).
  * factors: ||F - F_oracle||max / ||F_oracle||max <= 1e-10 (packed LU; Cholesky L).
"""
import json
import os

import numpy as np
import pytest

import rulinalg_amd as rl
from rulinalg_amd import _lib, workloads as wl
from oracle import oracle as orc
from tests.refcmp import assert_eq

pytestmark = pytest.mark.gpu

KATS = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "kats.json")))
U = np.finfo(np.float64).eps


@pytest.fixture(scope="module", autouse=True)
def lib():
    lib = _lib.load()
    assert lib.rlg_set_device(0) == 0, _lib.last_error()
    return lib


def eta(a, x, b):
    return np.abs(b - a @ x).max() / (np.abs(a).sum(1).max() * np.abs(x).max() + np.abs(b).max())


def fwd_tol(a):
    n = a.shape[0]
    k = np.linalg.cond(a, np.inf)
    return max(1e-13, 16 * n * k * U)


# ---------------------------------------------------------------- GEMM

@pytest.mark.parametrize("case", KATS["gemm"], ids=lambda c: c["name"])
def test_kat_gemm(case):
    dt = np.float32 if case.get("dtype") == "f32" else np.float64
    if "parent" in case:
        parent = rl.Matrix(np.array(case["parent"], dtype=dt))
        r0, c0, rows, cols = case["slice"]
        d = rl.MatrixSlice.from_matrix(parent, [r0, c0], rows, cols)
        c = d * d if case.get("b_is_slice") else d * rl.Matrix(np.array(case["b"], dtype=dt))
    else:
        c = rl.Matrix(np.array(case["a"], dtype=dt)) * rl.Matrix(np.array(case["b"], dtype=dt))
    assert_eq(c.to_numpy(), case["c"], case["comp"], dtype=dt, what=case["name"])


def test_gemm_shape_panic():
    with pytest.raises(rl.ShapePanic):
        rl.Matrix(np.ones((2, 3))) * rl.Matrix(np.ones((2, 3)))


@pytest.mark.parametrize("m,k,n", [(1, 1, 1), (3, 2, 3), (17, 23, 9), (128, 128, 128),
                                   (130, 257, 129), (256, 256, 256), (64, 1000, 33)])
def test_gemm_f64_vs_oracle(m, k, n):
    a = wl.uniform(m, k, 100 + m)
    b = wl.uniform(k, n, 200 + n)
    c = (rl.Matrix(a) * rl.Matrix(b)).to_numpy()
    c_or = orc.matmul(a, b)
    assert np.linalg.norm(c - c_or) / (np.linalg.norm(a) * np.linalg.norm(b)) <= 1e-13
    gamma = k * U / (1 - k * U)
    exact = a.astype(np.longdouble) @ b.astype(np.longdouble)
    assert np.all(np.abs(c - exact) <= 2 * gamma * (np.abs(a) @ np.abs(b)) + 1e-300)


def test_gemm_strided_operands():
    """MatrixSlice operands: row_stride = parent cols (slice.rs:15-33)."""
    pa = wl.uniform(300, 200, 7)
    pb = wl.uniform(260, 150, 8)
    A = rl.MatrixSlice.from_matrix(rl.Matrix(pa), [5, 3], 120, 140)
    B = rl.MatrixSlice.from_matrix(rl.Matrix(pb), [11, 2], 140, 97)
    c = (A * B).to_numpy()
    ref = orc.matmul(pa[5:125, 3:143], pb[11:151, 2:99])
    assert np.linalg.norm(c - ref) / (np.linalg.norm(pa[5:125, 3:143]) * np.linalg.norm(pb[11:151, 2:99])) <= 1e-13


def test_gemm_f32_vs_oracle():
    a = wl.uniform(200, 130, 3, np.float32)
    b = wl.uniform(130, 170, 4, np.float32)
    c = (rl.Matrix(a) * rl.Matrix(b)).to_numpy()
    c_or = orc.matmul(a, b)
    assert c.dtype == np.float32
    assert np.linalg.norm(c - c_or) / (np.linalg.norm(a) * np.linalg.norm(b)) <= 1e-6


@pytest.mark.parametrize("dt,m,k,n", [(np.float32, 131, 256, 260), (np.float32, 300, 36, 4),
                                      (np.float64, 131, 258, 130), (np.float64, 257, 2, 514)])
def test_gemm_vector_loads(dt, m, k, n):
    """Shapes whose rows are whole 16-byte vectors (k and n multiples of 4 in
    FP32, of 2 in FP64) take the kernel's 16-byte-load staging; m and n are
    not multiples of the 128 x 128 tile."""
    a = wl.uniform(m, k, 300 + m, dt)
    b = wl.uniform(k, n, 400 + n, dt)
    c = (rl.Matrix(a) * rl.Matrix(b)).to_numpy()
    c_or = orc.matmul(a, b)
    tol = 1e-6 if dt == np.float32 else 1e-13
    assert np.linalg.norm(c - c_or) / (np.linalg.norm(a) * np.linalg.norm(b)) <= tol


def test_gemm_dev_alpha_beta():
    lib = _lib.load()
    m, k, n = 150, 70, 90
    a, b, c0 = wl.uniform(m, k, 1), wl.uniform(k, n, 2), wl.uniform(m, n, 3)
    ptrs = []
    for arr in (a, b, c0):
        p = _lib.C.c_void_p()
        assert lib.rlg_dev_alloc(arr.nbytes, _lib.C.byref(p)) == 0
        assert lib.rlg_dev_upload(p, arr.ctypes.data, arr.nbytes) == 0
        ptrs.append(p)
    assert lib.rlg_gemm_f64_dev(m, k, n, -1.0, ptrs[0], k, ptrs[1], n, 1.0, ptrs[2], n, None) == 0
    out = np.empty_like(c0)
    assert lib.rlg_dev_download(out.ctypes.data, ptrs[2], out.nbytes) == 0
    for p in ptrs:
        lib.rlg_dev_free(p)
    assert np.abs(out - (c0 - a @ b)).max() < 1e-13


@pytest.mark.parametrize("case", KATS["matvec"], ids=lambda c: c["name"])
def test_kat_matvec(case):
    y = rl.Matrix(np.array(case["a"], float)) * rl.Vector(case["v"])
    assert_eq(y.data(), case["y"], case["comp"], what=case["name"])


def test_matvec_vs_oracle():
    a, v = wl.gaussian(300, 211, 5), wl.gaussian(211, 1, 6)[:, 0]
    y = (rl.Matrix(a) * rl.Vector(v)).data()
    assert np.abs(y - orc.matvec(a, v)).max() <= 1e-13 * np.abs(a).sum(1).max() * np.abs(v).max()


# ---------------------------------------------------------------- permutations

@pytest.mark.parametrize("name", ["left_mul_matrix", "right_mul_matrix"])
def test_kat_permutation_products(name):
    c = {c["name"]: c for c in KATS["permutation"]}[name]
    p = rl.PermutationMatrix(c["perm"])
    m = rl.Matrix(np.array(c["a"], float))
    out = (p * m) if name == "left_mul_matrix" else (m * p)
    assert out.to_numpy().tolist() == c["out"]


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("rows,cols", [(1, 1), (5, 3), (300, 777), (2049, 65)])
def test_permutation_products_vs_oracle(dt, rows, cols):
    """Row and column permutations on the device: bit-exact (data movement)."""
    r = np.random.default_rng(rows * 7 + cols)
    a = r.standard_normal((rows, cols)).astype(dt)
    pr = r.permutation(rows)
    pc = r.permutation(cols)
    assert np.array_equal((rl.PermutationMatrix(pr) * rl.Matrix(a)).to_numpy(), orc.permute_rows(pr, a))
    assert np.array_equal((rl.Matrix(a) * rl.PermutationMatrix(pc)).to_numpy(), orc.permute_cols(pc, a))


def test_permutation_product_errors():
    a = rl.Matrix(np.ones((3, 4)))
    with pytest.raises(rl.ShapePanic):
        rl.PermutationMatrix([0, 1]) * a          # perm size != rows
    with pytest.raises(rl.ShapePanic):
        a * rl.PermutationMatrix([0, 1, 2])       # perm size != cols
    with pytest.raises(rl.Error) as e:
        rl.PermutationMatrix([0, 0, 1]) * a       # not a permutation
    assert e.value.kind() == rl.ErrorKind.InvalidPermutation
    with pytest.raises(rl.Error) as e:
        a * rl.PermutationMatrix([0, 1, 2, 9])    # out of range
    assert e.value.kind() == rl.ErrorKind.InvalidPermutation


# ---------------------------------------------------------------- LU KATs

@pytest.mark.parametrize("case", KATS["lu_exact"], ids=lambda c: c["name"])
def test_kat_lu_exact(case):
    lup = rl.PartialPivLu.decompose(rl.Matrix(np.array(case["a"], float))).unpack()
    assert_eq(lup.l.to_numpy(), case["l"], case["comp"], what="L")
    assert_eq(lup.u.to_numpy(), case["u"], case["comp"], what="U")
    if "p_matrix" in case:
        assert_eq(lup.p.as_matrix().to_numpy(), case["p_matrix"], case["comp"], what="P")


@pytest.mark.parametrize("case", KATS["lu_reconstruct"], ids=lambda c: c["name"])
def test_kat_lu_reconstruct(case):
    a = np.array(case["a"], float)
    lup = rl.PartialPivLu.decompose(rl.Matrix(a)).unpack()
    k = lup.p.inverse() * (lup.l * lup.u)
    assert_eq(k.to_numpy(), a, case["comp"], what=case["name"])


@pytest.mark.parametrize("case", KATS["lu_inverse"], ids=lambda c: c["name"])
def test_kat_lu_inverse(case):
    inv = rl.PartialPivLu.decompose(rl.Matrix(np.array(case["a"], float))).inverse()
    assert_eq(inv.to_numpy(), case["inv"], case["comp"], what=case["name"])


@pytest.mark.parametrize("case", KATS["lu_det"], ids=lambda c: c["name"])
def test_kat_lu_det(case):
    d = rl.PartialPivLu.decompose(rl.Matrix(np.array(case["a"], float))).det()
    assert_eq(d, case["det"], case["comp"], what=case["name"])


@pytest.mark.parametrize("case", KATS["lu_solve"], ids=lambda c: c["name"])
def test_kat_lu_solve(case):
    x = rl.PartialPivLu.decompose(rl.Matrix(np.array(case["a"], float))).solve(rl.Vector(case["b"]))
    assert_eq(x.data(), case["x"], case["comp"], case.get("tol"), what=case["name"])


def test_kat_matrix_solve():
    """Matrix::solve (impl_mat.rs:331-343): [[2,3],[1,2]] x = [13,8] -> [2,3] exactly."""
    x = rl.Matrix(np.array([[2.0, 3.0], [1.0, 2.0]])).solve(rl.Vector([13.0, 8.0]))
    assert list(x.data()) == [2.0, 3.0]


@pytest.mark.parametrize("case", KATS["lu_singular"], ids=lambda c: c["name"])
def test_kat_lu_singular(case):
    with pytest.raises(rl.Error) as e:
        rl.PartialPivLu.decompose(rl.Matrix(np.array(case["a"], float)))
    assert e.value.kind() == rl.ErrorKind.DivByZero


def test_lu_nonsquare_panics():
    with pytest.raises(rl.ShapePanic):
        rl.PartialPivLu.decompose(rl.Matrix(np.ones((2, 3))))


def test_lu_solve_wrong_size_panics():
    lu = rl.PartialPivLu.decompose(rl.Matrix(np.eye(3)))
    with pytest.raises(rl.ShapePanic):
        lu.solve(rl.Vector([1.0, 2.0]))


def test_lu_empty():
    lu = rl.PartialPivLu.decompose(rl.Matrix(np.zeros((0, 0))))
    assert lu.det() == 1.0
    assert lu.solve(rl.Vector(np.zeros(0))).size() == 0


# ---------------------------------------------------------------- LU vs oracle

@pytest.mark.parametrize("n", [1, 2, 5, 31, 63, 64])
def test_lu_bit_exact_single_panel(n):
    """Orders <= 64 run entirely inside one panel with the reference's
    operation order: factor and pivots are bit-identical to the oracle."""
    a = wl.gaussian(n, n, 1000 + n)
    lu_or, perm_or, ipiv_or = orc.lu_decompose(a)
    lu = rl.PartialPivLu.decompose(rl.Matrix(a))
    assert np.array_equal(lu.ipiv(), ipiv_or)
    assert np.array_equal(lu.perm(), perm_or)
    assert np.array_equal(lu.packed(), lu_or)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n", [33, 48, 64])
def test_lu_single_panel_bits_with_zeros(n, dtype):
    """Small-integer matrices (exact ties and many exact zeros, so zero
    multipliers of either sign): inside one panel the packed factor equals
    the oracle's bit for bit, signs of zeros included.  The reference's
    elimination touches only the columns right of the pivot column
    (lu.rs:603-616), so a -0 multiplier stays -0; np.array_equal alone
    would accept +0 for it."""
    a = np.random.Generator(np.random.Philox(4000 + n)).integers(-3, 4, (n, n)).astype(dtype)
    lu_or, perm_or, ipiv_or = orc.lu_decompose(a, dtype=dtype)
    lu = rl.PartialPivLu.decompose(rl.Matrix(a))
    assert np.array_equal(lu.ipiv(), ipiv_or)
    got = np.ascontiguousarray(lu.packed(), dtype=dtype)
    want = np.ascontiguousarray(lu_or, dtype=dtype)
    bits = np.uint64 if dtype == np.float64 else np.uint32
    assert ((want == 0) & np.signbit(want)).any(), "the case must hold a -0"
    assert np.array_equal(got.view(bits), want.view(bits)), \
        f"{int((got.view(bits) != want.view(bits)).sum())} elements differ in their bits"


@pytest.mark.parametrize("n", [65, 127, 128, 129, 200, 255, 256, 257, 300, 513, 777, 2304])
def test_lu_vs_oracle(n):
    a = wl.gaussian(n, n, 2000 + n)
    lu_or, perm_or, ipiv_or, margin = orc.lu_decompose(a, with_margin=True)
    lu = rl.PartialPivLu.decompose(rl.Matrix(a))
    ipiv = lu.ipiv()
    clear = margin > 1e-9 * np.abs(a).max()
    first_unclear = np.argmin(clear) if not np.all(clear) else n
    # identical pivot sequence up to the first near-tie (afterwards rows may legitimately differ)
    assert np.array_equal(ipiv[:first_unclear], ipiv_or[:first_unclear])
    if first_unclear == n:
        assert np.array_equal(lu.perm(), perm_or)
        f = lu.packed()
        assert np.abs(f - lu_or).max() / np.abs(lu_or).max() <= 1e-10
    b = wl.gaussian(n, 1, 3000 + n)[:, 0]
    x = lu.solve(rl.Vector(b)).data()
    assert eta(a, x, b) <= 1e-13
    x_or = orc.lu_solve(lu_or, perm_or, b)
    assert np.abs(x - x_or).max() / np.abs(x_or).max() <= fwd_tol(a)


def test_lu_f32_and_mixed_wide_panels():
    """Order 4608: the first FP32 panels span nine 512-row workgroups that
    exchange their pivot candidates (FP64 panels span nine from 2304, above)."""
    n = 4608
    a = wl.gaussian(n, n, 4608)
    b = wl.gaussian(1, n, 4609)[0]
    lu = rl.PartialPivLu.decompose(rl.Matrix(a.astype(np.float32)))
    x = lu.solve(rl.Vector(b.astype(np.float32))).data().astype(np.float64)
    assert eta(a, x, b) <= 2e-5
    xm, rep = rl.solve_mixed(a, b, rl.MixedOptions(factor_prec=_lib.PREC_F32))
    assert rep["converged"] == 1 and rep["fell_back"] == 0
    assert eta(a, xm.reshape(-1), b) <= 1e-13


def test_lu_multi_rhs():
    n, r = 300, 7
    a = wl.gaussian(n, n, 41)
    b = wl.gaussian(r, n, 42)
    lu = rl.PartialPivLu.decompose(rl.Matrix(a))
    x = lu.solve_many(b)
    lu_or, perm_or, _ = orc.lu_decompose(a)
    for i in range(r):
        assert eta(a, x[i], b[i]) <= 1e-13
        x_or = orc.lu_solve(lu_or, perm_or, b[i])
        assert np.abs(x[i] - x_or).max() / np.abs(x_or).max() <= fwd_tol(a)


def test_lu_tie_breaking_across_workgroups():
    """An exact tie between rows 50 and 200, which live in different panel
    workgroups (128 rows each): the lower row index must win (lu.rs:170-178)."""
    n = 300
    a = wl.uniform(n, n, 31)
    a[:, 0] *= 0.5
    a[50, 0], a[200, 0] = 5.0, -5.0
    lu_or, perm_or, ipiv_or, margin = orc.lu_decompose(a, with_margin=True)
    assert ipiv_or[0] == 50
    lu = rl.PartialPivLu.decompose(rl.Matrix(a))
    ipiv = lu.ipiv()
    assert ipiv[0] == 50
    clear = margin[1:] > 1e-9
    first_unclear = 1 + (np.argmin(clear) if not np.all(clear) else n - 1)
    assert np.array_equal(ipiv[:first_unclear], ipiv_or[:first_unclear])


def test_lu_singular_large():
    """A structurally zero column (index 150, beyond the first panels) gives an
    exactly zero pivot under any update order, as in the reference."""
    n = 200
    a = wl.gaussian(n, n, 77)
    a[:, 150] = 0.0
    with pytest.raises(rl.Error) as e:
        rl.PartialPivLu.decompose(rl.Matrix(a))
    assert e.value.kind() == rl.ErrorKind.DivByZero


def test_lu_f32():
    n = 150
    a = wl.gaussian(n, n, 9, np.float32)
    lu_or, perm_or, ipiv_or = orc.lu_decompose(a, dtype=np.float32)
    lu = rl.PartialPivLu.decompose(rl.Matrix(a))
    assert lu.packed().dtype == np.float32
    b = wl.gaussian(n, 1, 10, np.float32)[:, 0]
    x = lu.solve(rl.Vector(b)).data()
    x_or = orc.lu_solve(lu_or, perm_or, b)
    assert np.abs(x - x_or).max() / np.abs(x_or).max() < 1e-3


@pytest.mark.parametrize("n", [40, 90])
def test_lu_det_inverse_vs_oracle(n):
    a = wl.gaussian(n, n, 55)
    lu_or, perm_or, _ = orc.lu_decompose(a)
    lu = rl.PartialPivLu.decompose(rl.Matrix(a))
    d, d_or = lu.det(), orc.lu_det(lu_or, perm_or)
    if n <= 64:
        assert d == d_or  # one panel: same factor bits, same fold (lu.rs:291-300)
    else:
        assert abs(d - d_or) <= 1e-12 * n * abs(d_or)
    inv = lu.inverse().to_numpy()
    assert np.abs(inv @ a - np.eye(n)).max() < 1e-11


def test_order256_parity():
    """The order-256 workload on the GPU: GEMM 256^3 U[-1,1] (seeds 0, 1), LU
    of a 256x256 N(0,1) (seed 2), one RHS (seed 3)."""
    c1 = wl.ORDER256
    n = c1["n"]
    a = wl.uniform(n, n, c1["gemm_a"])
    b = wl.uniform(n, n, c1["gemm_b"])
    c = (rl.Matrix(a) * rl.Matrix(b)).to_numpy()
    assert np.linalg.norm(c - orc.matmul(a, b)) / (np.linalg.norm(a) * np.linalg.norm(b)) <= 1e-13
    al = wl.gaussian(n, n, c1["lu_a"])
    bl = wl.gaussian(n, 1, c1["lu_b"])[:, 0]
    lu_or, perm_or, ipiv_or, margin = orc.lu_decompose(al, with_margin=True)
    lu = rl.PartialPivLu.decompose(rl.Matrix(al))
    assert np.all(margin > 1e-9)
    assert np.array_equal(lu.ipiv(), ipiv_or)
    x = lu.solve(rl.Vector(bl)).data()
    x_or = orc.lu_solve(lu_or, perm_or, bl)
    assert np.abs(x - x_or).max() / np.abs(x_or).max() <= 1e-13
    assert eta(al, x, bl) <= 1e-13


# ---------------------------------------------------------------- Cholesky

@pytest.mark.parametrize("case", KATS["chol_exact"], ids=lambda c: c["name"])
def test_kat_chol(case):
    l = rl.Cholesky.decompose(rl.Matrix(np.array(case["a"], float))).unpack()
    assert_eq(l.to_numpy(), case["l"], case["comp"], what=case["name"])


@pytest.mark.parametrize("case", KATS["chol_fail"], ids=lambda c: c["name"])
def test_kat_chol_fail(case):
    with pytest.raises(rl.Error) as e:
        rl.Cholesky.decompose(rl.Matrix(np.array(case["a"], float)))
    assert e.value.kind() == rl.ErrorKind.DecompFailure


@pytest.mark.parametrize("case", KATS["chol_det"], ids=lambda c: c["name"])
def test_kat_chol_det(case):
    d = rl.Cholesky.decompose(rl.Matrix(np.array(case["a"], float))).det()
    assert_eq(d, case["det"], case["comp"], what=case["name"])


@pytest.mark.parametrize("case", KATS["chol_solve"], ids=lambda c: c["name"])
def test_kat_chol_solve(case):
    x = rl.Cholesky.decompose(rl.Matrix(np.array(case["a"], float))).solve(rl.Vector(case["b"]))
    assert_eq(x.data(), case["x"], case["comp"], what=case["name"])


@pytest.mark.parametrize("case", KATS["chol_inverse"], ids=lambda c: c["name"])
def test_kat_chol_inverse(case):
    inv = rl.Cholesky.decompose(rl.Matrix(np.array(case["a"], float))).inverse()
    assert_eq(inv.to_numpy(), case["inv"], case["comp"], what=case["name"])


def test_chol_identity_property():
    for n in [0, 1, 2, 7, 30]:
        l = rl.Cholesky.decompose(rl.Matrix.identity(n)).unpack()
        assert_eq(l.to_numpy(), np.eye(n), "float")


@pytest.mark.parametrize("n", [1, 3, 17, 64])
def test_chol_bit_exact_single_block(n):
    a = wl.spd(n, 500 + n)
    w_or = orc.chol_decompose(a)
    l = rl.Cholesky.decompose(rl.Matrix(a)).unpack().to_numpy()
    assert np.array_equal(l, orc.chol_unpack(w_or))


@pytest.mark.parametrize("n", [33, 64])
def test_chol_single_block_bits_banded(n):
    """A pentadiagonal integer SPD matrix (4 on the diagonal, -1 on the two
    sub- and superdiagonals each side): its factor holds exact zeros outside
    the band; inside one block the factor equals the oracle's bit for bit,
    signs of zeros included (np.array_equal alone would accept either)."""
    a = 6.0 * np.eye(n)
    for d in (1, 2):
        a -= np.eye(n, k=d) + np.eye(n, k=-d)
    l_or = orc.chol_unpack(orc.chol_decompose(a))
    l = rl.Cholesky.decompose(rl.Matrix(a)).unpack().to_numpy()
    assert (l_or == 0).sum() > n * (n - 5) // 2, "the factor must hold exact zeros"
    assert np.array_equal(l.view(np.uint64), l_or.view(np.uint64))


@pytest.mark.parametrize("n", [1, 40, 64, 150, 520])
def test_chol_f32_vs_oracle(n):
    """Cholesky<f32> through rlg_chol_decompose_f32, rlg_chol_unpack_f32 and
    rlg_chol_solve_f32: bit-identical to the oracle's FP32 restatement inside
    one 64-column block, within FP32 rounding beyond it
    (||L - L_oracle||max / ||L_oracle||max <= 1e-5); the solve's normwise
    backward error, in FP64 arithmetic on the FP32 data, <= 64 n u_f32."""
    a = wl.spd(n, 700 + n).astype(np.float32)
    w_or = orc.chol_decompose(a, dtype=np.float32)
    l_or = orc.chol_unpack(w_or)
    ch = rl.Cholesky.decompose(rl.Matrix(a))
    l = ch.unpack().to_numpy()
    assert l.dtype == np.float32
    if n <= 64:
        assert np.array_equal(l.view(np.uint32), l_or.view(np.uint32))
    else:
        assert np.abs(l - l_or).max() / np.abs(l_or).max() <= 1e-5
    b = wl.gaussian(n, 1, 750 + n)[:, 0].astype(np.float32)
    x = ch.solve(rl.Vector(b, dtype=np.float32)).data()
    assert x.dtype == np.float32
    u32 = np.finfo(np.float32).eps
    assert eta(a.astype(np.float64), x.astype(np.float64), b.astype(np.float64)) <= 64 * n * u32
    x_or = orc.chol_solve(w_or, b)
    assert np.abs(x - x_or).max() / np.abs(x_or).max() <= 1e-4


@pytest.mark.parametrize("n", [1, 37, 64, 300])
def test_lu_unpack_entry_point(n):
    """rlg_lu_unpack_f64 (PartialPivLu::unpack, lu.rs:135-150) called directly:
    L with its unit diagonal, U and the stored permutation equal the oracle's
    unpack of the oracle's factor, bit for bit inside one panel (n <= 64) and
    within 1e-10 of the largest entry beyond it, and P A = L U."""
    a = wl.gaussian(n, n, 5000 + n)
    lu = rl.PartialPivLu.decompose(rl.Matrix(a))
    l = np.empty((n, n))
    u = np.empty((n, n))
    perm = np.empty(n, dtype=np.uintp)
    st = _lib.load().rlg_lu_unpack_f64(lu._h, l.ctypes.data, u.ctypes.data, perm.ctypes.data)
    assert st == 0, _lib.last_error()
    lu_or, perm_or, _ = orc.lu_decompose(a)
    l_or, u_or = orc.lu_unpack(lu_or)
    assert np.array_equal(perm, perm_or)
    if n <= 64:
        assert np.array_equal(l.view(np.uint64), l_or.view(np.uint64))
        assert np.array_equal(u.view(np.uint64), u_or.view(np.uint64))
    else:
        assert np.abs(l - l_or).max() <= 1e-10 and np.abs(u - u_or).max() / np.abs(u_or).max() <= 1e-10
    assert np.allclose(np.diag(l), 1.0) and np.all(np.triu(l, 1) == 0) and np.all(np.tril(u, -1) == 0)
    rec = orc.permute_rows(orc.perm_inverse(perm), l @ u)
    assert np.abs(rec - a).max() <= 1e-12 * np.abs(a).max() * n


@pytest.mark.parametrize("n", [65, 128, 200, 256, 257, 520])
def test_chol_vs_oracle(n):
    a = wl.spd(n, 600 + n)
    junk = a.copy()
    junk[np.triu_indices(n, 1)] = 1e6  # only the lower triangle may be read
    w_or = orc.chol_decompose(a)
    l_or = orc.chol_unpack(w_or)
    ch = rl.Cholesky.decompose(rl.Matrix(junk))
    l = ch.unpack().to_numpy()
    assert np.abs(l - l_or).max() / np.abs(l_or).max() <= 1e-10
    b = wl.gaussian(n, 1, 700 + n)[:, 0]
    x = ch.solve(rl.Vector(b)).data()
    assert eta(a, x, b) <= 1e-13
    x_or = orc.chol_solve(w_or, b)
    assert np.abs(x - x_or).max() / np.abs(x_or).max() <= fwd_tol(a)


def test_chol_indefinite_large():
    n = 150
    a = wl.spd(n, 9)
    a[100, 100] = -5.0
    with pytest.raises(rl.Error) as e:
        rl.Cholesky.decompose(rl.Matrix(a))
    assert e.value.kind() == rl.ErrorKind.DecompFailure


def test_chol_nonsquare_panics():
    with pytest.raises(rl.ShapePanic):
        rl.Cholesky.decompose(rl.Matrix(np.ones((3, 2))))


# ---------------------------------------------------------------- mixed

@pytest.mark.parametrize("prec", [_lib.PREC_F32, _lib.PREC_F16, _lib.PREC_BF16])
def test_mixed_refined_well_conditioned(prec):
    """kappa = 1e2 at order 1600 (three depth-512 low-precision trailing
    updates, 1600 = 3 * 512 + 64): kappa * u is 1e-5 (FP32), 0.05 (FP16) and
    0.4 (BF16), so refinement must converge without the FP64 fallback in every
    precision, which also shows the low-precision updates are right."""
    n = 1600
    a = wl.conditioned(n, 1e2, 1, 2)
    xt = wl.gaussian(n, 2, 3).T.copy()
    b = xt @ a.T
    x, rep = rl.solve_mixed(a, b, rl.MixedOptions(factor_prec=prec))
    assert rep["converged"] == 1 and rep["fell_back"] == 0, rep
    for i in range(2):
        assert eta(a, x[i], b[i]) <= 1e-13


@pytest.mark.parametrize("pivot", [_lib.PIVOT_TOURNAMENT, _lib.PIVOT_PARTIAL], ids=["tournament", "partial"])
@pytest.mark.parametrize("prec,n", [(_lib.PREC_F32, 200), (_lib.PREC_F32, 1500), (_lib.PREC_F32, 4100),
                                    (_lib.PREC_F16, 200), (_lib.PREC_F16, 1500),
                                    (_lib.PREC_F16X3, 1500), (_lib.PREC_F16X3, 4100)])
def test_mixed_pivoting_modes(prec, n, pivot):
    """TSLU panels (256-row leaves merged four at a time: n = 1500 is a
    three-level tree, n = 4100 a four-level one whose last panel is 4 columns
    wide) and the exchange panel of partial pivoting: refinement
    must converge without the FP64 fallback in both modes."""
    a = wl.conditioned(n, 1e3, 5, 6)
    b = wl.gaussian(2, n, 7)
    x, rep = rl.solve_mixed(a, b, rl.MixedOptions(factor_prec=prec, pivot=pivot))
    assert rep["converged"] == 1 and rep["fell_back"] == 0
    for i in range(2):
        assert eta(a, x[i], b[i]) <= 1e-13
    with pytest.raises(rl.Error) as e:  # an unknown pivoting mode is an argument error
        rl.solve_mixed(a, b, rl.MixedOptions(pivot=7))
    assert e.value.kind() == rl.ErrorKind.InvalidArg


def test_mixed_kappa_1e4_f32():
    n = 512
    a = wl.conditioned(n, 1e4, 4, 5)
    b = wl.gaussian(n, 1, 6)[:, 0]
    x, rep = rl.solve_mixed(a, b, rl.MixedOptions(factor_prec=_lib.PREC_F32))
    assert rep["converged"] == 1 and rep["fell_back"] == 0
    assert eta(a, x, b) <= 1e-13
    lu_or, perm_or, _ = orc.lu_decompose(a)
    x_or = orc.lu_solve(lu_or, perm_or, b)
    assert np.abs(x - x_or).max() / np.abs(x_or).max() <= fwd_tol(a)


def test_mixed_fallback_on_stall():
    """kappa = 1e12 is far beyond what FP32 refinement can recover: the solver
    must notice the stall and fall back to the plain FP64 factorization."""
    n = 300
    a = wl.conditioned(n, 1e12, 7, 8)
    b = wl.gaussian(n, 1, 9)[:, 0]
    x, rep = rl.solve_mixed(a, b, rl.MixedOptions(factor_prec=_lib.PREC_F32))
    assert rep["fell_back"] == 1
    assert eta(a, x, b) <= 1e-13
    with pytest.raises(Exception):
        rl.solve_mixed(a, b, rl.MixedOptions(factor_prec=_lib.PREC_F32, fallback=False))
