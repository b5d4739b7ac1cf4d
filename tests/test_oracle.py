"""Pins the CPU oracle (oracle/rlg_oracle.c) before anything is checked against it.

Checks, per oracle function:
  * the reference's own known-answer tests, transcribed in tests/golden/kats.json,
    under the reference's comparators;
  * LAPACK (scipy.linalg.lu_factor = dgetrf, same first-maximum pivot rule via
    idamax) for the pivot sequence and the packed factors on seeded Gaussian
    matrices, and numpy's Cholesky (dpotrf) for the Cholesky factor;
  * hand-computed cases whose answer depends on the reference's exact operation
    order (utils::dot's eight-lane order, the forward substitution's sequential
    fold), so that a plausible restatement mistake fails;
  * identities: P^-1 L U = A, L L^T = A, backward error of solves, det = product
    of pivots times parity, inverse * A = I.
"""
import json
import math
import os

import numpy as np
import pytest

from oracle import oracle as orc
from tests.refcmp import assert_eq, compare

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "kats.json")
KATS = json.load(open(GOLDEN))


def rng(seed):
    return np.random.Generator(np.random.Philox(seed))


# ---------------------------------------------------------------- KATs: GEMM

@pytest.mark.parametrize("case", KATS["gemm"], ids=lambda c: c["name"])
def test_kat_gemm(case):
    dt = np.float32 if case.get("dtype") == "f32" else np.float64
    if "parent" in case:
        parent = np.array(case["parent"], dtype=dt)
        r0, c0, rows, cols = case["slice"]
        ld = parent.shape[1]
        flat = parent.ravel()
        a_view = flat[r0 * ld + c0:]
        if case.get("b_is_slice"):
            c = orc.matmul(a_view, a_view, lda=ld, ldb=ld, rows=rows, inner=cols, cols=cols)
        else:
            b = np.array(case["b"], dtype=dt)
            c = orc.matmul(a_view, b.ravel(), lda=ld, ldb=b.shape[1], rows=rows, inner=cols,
                           cols=b.shape[1])
    else:
        c = orc.matmul(np.array(case["a"], dtype=dt), np.array(case["b"], dtype=dt))
    assert_eq(c, case["c"], case["comp"], dtype=dt, what=case["name"])


def test_gemm_shape_mismatch_is_error():
    """mat_mul.rs:21 asserts self.cols == m.rows; the oracle reports it as a status."""
    assert orc.lib().orc_matmul_f64(2, 3, 2, 2, None, 3, None, 2, None) == orc.E_SHAPE


def test_gemm_against_compensated_product():
    """Elementwise bound |C - fl(C)| <= gamma_k |A||B| against a math.fsum product."""
    r = rng(11)
    a = r.uniform(-1, 1, (17, 23))
    b = r.uniform(-1, 1, (23, 9))
    c = orc.matmul(a, b)
    exact = np.array([[math.fsum(a[i, k] * b[k, j] for k in range(23)) for j in range(9)]
                      for i in range(17)])
    gamma = 23 * np.finfo(float).eps / (1 - 23 * np.finfo(float).eps)
    assert np.all(np.abs(c - exact) <= gamma * (np.abs(a) @ np.abs(b)))


def test_gemm_loop_order_is_ikj():
    """mat_mul.rs:76-98 accumulates over k in increasing order per (i, j):
    with a row [1e16, 1, -1e16] and ones, ((0 + 1e16) + 1) + -1e16 = 0 in binary64."""
    a = np.array([[1e16, 1.0, -1e16]])
    b = np.ones((3, 1))
    assert orc.matmul(a, b)[0, 0] == 0.0


# ------------------------------------------------------------ KATs: vectors

@pytest.mark.parametrize("case", KATS["matvec"], ids=lambda c: c["name"])
def test_kat_matvec(case):
    assert_eq(orc.matvec(np.array(case["a"], float), np.array(case["v"], float)), case["y"],
              case["comp"], what=case["name"])


def test_dot_eight_lane_order():
    """utils::dot combines partial sums as ((((s+p0)+p4)+p1)+p5)...; for the
    products [1e16, 1, 0, 0, -1e16, 1, 0, 0] that gives 2, whereas a
    sequential sum gives 1 and a pairwise sum gives 0."""
    u = np.array([1e16, 1.0, 0, 0, -1e16, 1.0, 0, 0])
    assert orc.dot(u, np.ones(8)) == 2.0
    # the tail after the 8-chunk is added sequentially: 2 + 1e16 = 1e16 + 2 (exact,
    # ulp is 2 there); + 1 = 1e16 + 3, a tie that rounds to even, 1e16 + 4; - 1e16 = 4
    u2 = np.concatenate([u, [1e16, 1.0, -1e16]])
    assert orc.dot(u2, np.ones(11)) == 4.0
    assert orc.dot(np.arange(1, 20, dtype=float), np.arange(1, 20, dtype=float)) == sum(
        i * i for i in range(1, 20))


def test_dot_f32():
    u = np.arange(1, 11, dtype=np.float32)
    assert orc.dot(u, u) == np.float32(385.0)


# ---------------------------------------------------------------- KATs: LU

@pytest.mark.parametrize("case", KATS["lu_exact"], ids=lambda c: c["name"])
def test_kat_lu_exact(case):
    lu, perm, ipiv = orc.lu_decompose(np.array(case["a"], float))
    l, u = orc.lu_unpack(lu)
    assert_eq(l, case["l"], case["comp"], what="L")
    assert_eq(u, case["u"], case["comp"], what="U")
    if "p_matrix" in case:
        assert_eq(orc.perm_matrix(perm), case["p_matrix"], case["comp"], what="P")


@pytest.mark.parametrize("case", KATS["lu_reconstruct"], ids=lambda c: c["name"])
def test_kat_lu_reconstruct(case):
    a = np.array(case["a"], float)
    lu, perm, _ = orc.lu_decompose(a)
    l, u = orc.lu_unpack(lu)
    # p.inverse() * l * u, computed with the oracle's own product
    lu_prod = orc.matmul(l, u)
    rec = orc.perm_matrix(orc.perm_inverse(perm)) @ lu_prod  # exact row permutation
    assert_eq(rec, a, case["comp"], what=case["name"])


@pytest.mark.parametrize("case", KATS["lu_inverse"], ids=lambda c: c["name"])
def test_kat_lu_inverse(case):
    lu, perm, _ = orc.lu_decompose(np.array(case["a"], float))
    assert_eq(orc.lu_inverse(lu, perm), case["inv"], case["comp"], what=case["name"])


@pytest.mark.parametrize("case", KATS["lu_det"], ids=lambda c: c["name"])
def test_kat_lu_det(case):
    lu, perm, _ = orc.lu_decompose(np.array(case["a"], float))
    assert_eq(orc.lu_det(lu, perm), case["det"], case["comp"], what=case["name"])


@pytest.mark.parametrize("case", KATS["lu_solve"], ids=lambda c: c["name"])
def test_kat_lu_solve(case):
    lu, perm, _ = orc.lu_decompose(np.array(case["a"], float))
    x = orc.lu_solve(lu, perm, np.array(case["b"], float))
    assert_eq(x, case["x"], case["comp"], case.get("tol"), what=case["name"])


@pytest.mark.parametrize("case", KATS["lu_singular"], ids=lambda c: c["name"])
def test_kat_lu_singular(case):
    with pytest.raises(orc.OracleError) as e:
        orc.lu_decompose(np.array(case["a"], float))
    assert e.value.code == orc.E_DIV_BY_ZERO


@pytest.mark.parametrize("case", KATS["lu_forward_substitution"], ids=lambda c: c["name"])
def test_kat_lu_forward_substitution(case):
    x = orc.lu_forward_substitution(np.array(case["lu"], float), np.array(case["b"], float))
    assert_eq(x, case["x"], case["comp"], what=case["name"])


def test_lu_forward_substitution_is_sequential_fold():
    """lu.rs:632-637 folds sequentially (not utils::dot): with an 8-long row
    [1e16, 1, 0, 0, -1e16, 1, 0, 0] and x = 1 the adjustment is 1, not 2."""
    n = 9
    lu = np.eye(n)
    lu[8, :8] = [1e16, 1.0, 0, 0, -1e16, 1.0, 0, 0]
    b = np.ones(n)
    x = orc.lu_forward_substitution(lu, b)
    assert x[8] == 0.0  # 1 - 1


def test_back_substitution_uses_dot_order():
    """back_substitution (mod.rs:347-353) uses utils::dot: row 0 of U has the
    eight off-diagonal entries [1e16, 1, 0, 0, -1e16, 1, 0, 0] against x = 1."""
    n = 9
    u = np.eye(n)
    u[0, 1:] = [1e16, 1.0, 0, 0, -1e16, 1.0, 0, 0]
    x = orc.back_substitution(u, np.ones(n))
    assert x[0] == 1.0 - 2.0


def test_lu_empty():
    lu, perm, ipiv = orc.lu_decompose(np.zeros((0, 0)))
    assert lu.shape == (0, 0) and len(perm) == 0
    assert orc.lu_det(lu, perm) == 1.0


def test_lu_nonsquare_is_shape_error():
    with pytest.raises(orc.OracleError) as e:
        orc.lu_decompose(np.ones((2, 3)))
    assert e.value.code == orc.E_SHAPE


@pytest.mark.parametrize("n,seed", [(8, 1), (33, 2), (64, 3), (200, 4), (256, 2)])
def test_lu_matches_lapack(n, seed):
    """Same pivot rule as LAPACK dgetrf (idamax: first index of max |.|), so the
    step pivots agree wherever the oracle's margins are clear of rounding."""
    sl = pytest.importorskip("scipy.linalg")
    a = rng(seed).standard_normal((n, n))
    lu, perm, ipiv, margin = orc.lu_decompose(a, with_margin=True)
    lu_l, piv_l = sl.lu_factor(a)
    scale = np.abs(a).max()
    clear = margin > 1e-10 * scale * n
    assert np.all(ipiv[clear] == piv_l[clear])
    if np.all(clear):
        assert np.array_equal(ipiv, piv_l)
        err = np.abs(lu - lu_l).max() / np.abs(lu_l).max()
        assert err < 1e-11 * n
    # stored permutation: y = p * b gives y[k] = b[raw[k]] with raw the swap product
    raw = np.arange(n)
    for k in range(n):
        raw[[k, ipiv[k]]] = raw[[ipiv[k], k]]
    assert np.array_equal(orc.perm_inverse(perm), raw)


@pytest.mark.parametrize("n", [1, 2, 7, 50, 130])
def test_lu_reconstruction_and_solve(n):
    r = rng(100 + n)
    a = r.standard_normal((n, n))
    lu, perm, _ = orc.lu_decompose(a)
    l, u = orc.lu_unpack(lu)
    raw = orc.perm_inverse(perm)
    rec = (l @ u)
    assert np.abs(rec - a[raw]).max() <= 8 * n * np.finfo(float).eps * np.abs(a).max() * max(1, np.abs(l).sum(1).max())
    b = r.standard_normal(n)
    x = orc.lu_solve(lu, perm, b)
    eta = np.abs(b - a @ x).max() / (np.abs(a).sum(1).max() * np.abs(x).max() + np.abs(b).max())
    assert eta < 10 * n * np.finfo(float).eps
    assert np.allclose(x, np.linalg.solve(a, b), rtol=1e-9, atol=1e-9)
    # det: product of pivots times parity (lu.rs:291-300) vs numpy's slogdet
    sign, logdet = np.linalg.slogdet(a)
    d = orc.lu_det(lu, perm)
    assert np.sign(d) == sign and abs(math.log(abs(d)) - logdet) < 1e-9 * max(1, abs(logdet))
    if n <= 50:
        inv = orc.lu_inverse(lu, perm)
        assert np.abs(inv @ a - np.eye(n)).max() < 1e-10 * n


def test_lu_f32():
    a = rng(5).standard_normal((40, 40)).astype(np.float32)
    lu, perm, ipiv = orc.lu_decompose(a, dtype=np.float32)
    assert lu.dtype == np.float32
    b = rng(6).standard_normal(40).astype(np.float32)
    x = orc.lu_solve(lu, perm, b)
    assert np.allclose(x, np.linalg.solve(a.astype(float), b.astype(float)), rtol=1e-3, atol=1e-3)
    # f32 epsilon governs the singularity test
    with pytest.raises(orc.OracleError):
        orc.lu_decompose(np.array([[1e-8, 0], [0, 1]], np.float32), dtype=np.float32)
    orc.lu_decompose(np.array([[1e-8, 0], [0, 1]]))  # fine in f64


def test_lu_singular_step_reported():
    a = np.array([[2.0, 1, 1], [4, 2, 2], [1, 5, 3]])
    # step 0 pivots on row 1 (4); row 0 becomes [0, 0, 0]; step 1 picks row 2; step 2 fails
    with pytest.raises(orc.OracleError) as e:
        orc.lu_decompose(a)
    assert e.value.code == orc.E_DIV_BY_ZERO and e.value.where == 2


# ---------------------------------------------------------------- Cholesky

@pytest.mark.parametrize("case", KATS["chol_exact"], ids=lambda c: c["name"])
def test_kat_chol(case):
    w = orc.chol_decompose(np.array(case["a"], float))
    assert_eq(orc.chol_unpack(w), case["l"], case["comp"], what=case["name"])


@pytest.mark.parametrize("case", KATS["chol_fail"], ids=lambda c: c["name"])
def test_kat_chol_fail(case):
    with pytest.raises(orc.OracleError) as e:
        orc.chol_decompose(np.array(case["a"], float))
    assert e.value.code == orc.E_DECOMP_FAILURE


@pytest.mark.parametrize("case", KATS["chol_det"], ids=lambda c: c["name"])
def test_kat_chol_det(case):
    w = orc.chol_decompose(np.array(case["a"], float))
    assert_eq(orc.chol_det(w), case["det"], case["comp"], what=case["name"])


@pytest.mark.parametrize("case", KATS["chol_solve"], ids=lambda c: c["name"])
def test_kat_chol_solve(case):
    w = orc.chol_decompose(np.array(case["a"], float))
    assert_eq(orc.chol_solve(w, np.array(case["b"], float)), case["x"], case["comp"], what=case["name"])


@pytest.mark.parametrize("case", KATS["chol_inverse"], ids=lambda c: c["name"])
def test_kat_chol_inverse(case):
    w = orc.chol_decompose(np.array(case["a"], float))
    assert_eq(orc.chol_inverse(w), case["inv"], case["comp"], what=case["name"])


@pytest.mark.parametrize("case", KATS["transpose_back_substitution"], ids=lambda c: c["name"])
def test_kat_transpose_back_substitution(case):
    x = orc.transpose_back_substitution(np.array(case["l"], float), np.array(case["b"], float))
    assert_eq(x, case["x"], case["comp"], what=case["name"])


@pytest.mark.parametrize("case", KATS["triangular"], ids=lambda c: c["name"])
def test_kat_triangular(case):
    m = np.array(case["m"], float)
    fn = orc.back_substitution if case["kind"] == "upper" else orc.forward_substitution
    if "err" in case:
        with pytest.raises(orc.OracleError) as e:
            fn(m, np.array(case["b"], float))
        assert e.value.code == orc.E_DIV_BY_ZERO
    else:
        assert_eq(fn(m, np.array(case["b"], float)), case["x"], case["comp"], what=case["name"])


def test_chol_identity_property():
    """property_cholesky_of_identity_is_identity (cholesky.rs:539-555), n <= 30."""
    for n in range(0, 31):
        w = orc.chol_decompose(np.eye(n))
        assert_eq(orc.chol_unpack(w), np.eye(n), "float")


def test_chol_reads_lower_triangle_only():
    r = rng(21)
    g = r.standard_normal((40, 40))
    a = g @ g.T / 40 + np.eye(40)
    junk = a.copy()
    junk[np.triu_indices(40, 1)] = r.standard_normal(40 * 39 // 2) * 1e3
    w1 = orc.chol_decompose(a)
    w2 = orc.chol_decompose(junk)
    assert np.array_equal(orc.chol_unpack(w1), orc.chol_unpack(w2))
    # the strict upper triangle still holds the input (unpack nullifies it)
    assert np.array_equal(np.triu(w2, 1), np.triu(junk, 1))


@pytest.mark.parametrize("n", [1, 5, 64, 150])
def test_chol_matches_lapack(n):
    r = rng(300 + n)
    g = r.standard_normal((n, n))
    a = g @ g.T / n + np.eye(n)
    w = orc.chol_decompose(a)
    l = orc.chol_unpack(w)
    l_ref = np.linalg.cholesky(a)
    assert np.abs(l - l_ref).max() < 1e-13 * n
    b = r.standard_normal(n)
    x = orc.chol_solve(w, b)
    assert np.allclose(x, np.linalg.solve(a, b), rtol=1e-10, atol=1e-10)
    assert abs(orc.chol_det(w) / np.linalg.det(a) - 1) < 1e-10 * n


def test_chol_indefinite_fails():
    with pytest.raises(orc.OracleError) as e:
        orc.chol_decompose(np.array([[1.0, 2], [2, 1]]))
    assert e.value.code == orc.E_DECOMP_FAILURE and e.value.where == 1


# ------------------------------------------------------------ permutations

def test_kat_permutations():
    cases = {c["name"]: c for c in KATS["permutation"]}
    c = cases["swap_rows"]
    raw = list(range(c["n"]))
    for i, j in c["swaps"]:  # PermutationMatrix::swap_rows = perm.swap(i, j)
        raw[i], raw[j] = raw[j], raw[i]
    assert raw == c["perm"]
    assert list(orc.perm_inverse(cases["inverse"]["perm"])) == cases["inverse"]["inv"]
    pv = cases["permute_vector"]
    y = orc.perm_matrix(pv["perm"]) @ np.array(pv["v"], float)
    assert list(y) == pv["y"]
    for cc in cases["parity"]["cases"]:
        assert orc.perm_parity(cc["perm"]) == cc["parity"]
    for cc in cases["det"]["cases"]:
        assert orc.perm_parity(cc["perm"]) == cc["det"]


def test_kat_permutation_products():
    """P * M and M * P (impl_permutation_mul.rs tests), and both against the
    full-matrix products they equal: P * M = as_matrix(P) @ M, and M * P =
    M @ as_matrix(P)^T (the reference's M * P scatters columns to perm[j])."""
    cases = {c["name"]: c for c in KATS["permutation"]}
    for name, fn in (("left_mul_matrix", orc.permute_rows), ("right_mul_matrix", orc.permute_cols)):
        c = cases[name]
        assert fn(c["perm"], np.array(c["a"], float)).tolist() == c["out"]
    r = np.random.default_rng(3)
    for n, m in ((1, 4), (7, 5), (30, 11)):
        p = r.permutation(n)
        a = r.standard_normal((n, m))
        assert np.array_equal(orc.permute_rows(p, a), orc.perm_matrix(p) @ a)
        b = r.standard_normal((m, n))
        assert np.array_equal(orc.permute_cols(p, b), b @ orc.perm_matrix(p).T)


def test_perm_parity_brute_force():
    """Parity from permute_by_swap equals (-1)^(inversions) on random permutations."""
    r = rng(9)
    for n in range(0, 9):
        for _ in range(5):
            p = r.permutation(n)
            inv = sum(1 for i in range(n) for j in range(i + 1, n) if p[i] > p[j])
            assert orc.perm_parity(p) == (1 if inv % 2 == 0 else -1)


def test_lu_solve_applies_stored_permutation():
    """solve computes y[perm[i]] = b[i] (permutation_matrix.rs:369-382) with
    perm the inverse of the swap product, i.e. y = P b with (P b)[k] = b[raw[k]]."""
    a = np.array([[0.0, 1.0], [1.0, 0.0]])
    lu, perm, ipiv = orc.lu_decompose(a)
    assert list(ipiv) == [1, 1] and list(perm) == [1, 0]
    assert list(orc.lu_solve(lu, perm, np.array([3.0, 5.0]))) == [5.0, 3.0]


def test_gemm_blocked_standin():
    r = rng(12)
    a = r.uniform(-1, 1, (70, 300))
    b = r.uniform(-1, 1, (300, 530))
    c = orc.gemm_blocked(a, b)
    assert np.abs(c - a @ b).max() < 1e-12


# ---------------------------------------------------------------- FullPivLu

def _fplu_reconstruct(lu, p, q):
    """p.inverse() * L * U * q.inverse() from the packed factor and the stored
    permutations with the reference's own permutation products
    (impl_permutation_mul.rs), as its reconstruction test does (lu.rs:903-907)."""
    l, u = orc.lu_unpack(lu)
    return orc.permute_cols(orc.perm_inverse(q), orc.permute_rows(orc.perm_inverse(p), l @ u))


@pytest.mark.parametrize("case", KATS["fplu"], ids=lambda c: c["name"])
def test_kat_fplu(case):
    a = np.array(case["a"], float).reshape(len(case["a"]), -1) if case["a"] else np.zeros((0, 0))
    lu, p, q, prow, pcol, steps = orc.fplu_decompose(a)
    if case.get("reconstruct"):
        assert_eq(_fplu_reconstruct(lu, p, q), a, case["comp"], what=case["name"])
        l, u = orc.lu_unpack(lu)
        assert np.array_equal(np.tril(l), l) and np.array_equal(np.triu(u), u)
    if "rank" in case:
        assert orc.fplu_rank(lu) == case["rank"]
    if "invertible" in case:
        assert orc.fplu_is_invertible(lu) == case["invertible"]
    if "det" in case:
        assert orc.fplu_det(lu, p, q) == case["det"]
    if "x" in case:
        x = orc.fplu_solve(lu, p, q, np.array(case["b"], float))
        assert_eq(x, case["x"], case["comp"], tol=case.get("tol"), what=case["name"])
    if "inv" in case:
        assert_eq(orc.fplu_inverse(lu, p, q), case["inv"], case["comp"], what=case["name"])
    if case.get("inverse_err") == "DivByZero":
        with pytest.raises(orc.OracleError) as e:
            orc.fplu_inverse(lu, p, q)
        assert e.value.code == orc.E_DIV_BY_ZERO
    if a.size == 0:
        assert orc.fplu_inverse(lu, p, q).shape == (0, 0)


@pytest.mark.parametrize("n,seed", [(5, 1), (17, 2), (64, 3), (150, 4)])
def test_fplu_matches_lapack_getc2(n, seed):
    """LAPACK dgetc2 is LU with complete pivoting over the same row-major scan;
    on Gaussian matrices (no ties, no tiny pivots) its row and column
    interchanges must equal the oracle's step pivots and its factor the
    oracle's packed factor to rounding."""
    lap = pytest.importorskip("scipy.linalg.lapack")
    a = rng(seed).standard_normal((n, n))
    lu, p, q, prow, pcol, steps = orc.fplu_decompose(a)
    lu_l, ipiv, jpiv, info = lap.dgetc2(a)
    assert info == 0 and steps == n
    assert np.array_equal(prow, ipiv) and np.array_equal(pcol, jpiv)  # scipy returns 0-based
    assert np.abs(lu - lu_l).max() / np.abs(lu_l).max() < 1e-12 * n


def test_fplu_complete_pivoting_bounds():
    """Properties only complete pivoting has: every multiplier |l_ij| <= 1 and
    every U row is dominated by its diagonal (|u_kj| <= |u_kk|), on a matrix
    where partial (column-only) pivoting violates the second."""
    a = np.array([[1.0, 100.0, 0.0], [2.0, 1.0, 3.0], [0.5, 4.0, 1.0]])
    lu, p, q, prow, pcol, steps = orc.fplu_decompose(a)
    l, u = orc.lu_unpack(lu)
    assert np.abs(np.tril(l, -1)).max() <= 1.0
    assert np.all(np.abs(u) <= np.abs(np.diag(u))[:, None] + 0.0)
    lu_pp, _, _ = orc.lu_decompose(a)
    u_pp = np.triu(lu_pp)
    assert not np.all(np.abs(u_pp) <= np.abs(np.diag(u_pp))[:, None])
    r = rng(13)
    for n in (6, 40):
        b = r.standard_normal((n, n))
        lu, p, q, *_ = orc.fplu_decompose(b)
        l, u = orc.lu_unpack(lu)
        assert np.abs(np.tril(l, -1)).max() <= 1.0
        assert np.all(np.abs(u) <= np.abs(np.diag(u))[:, None])
        assert np.abs(_fplu_reconstruct(lu, p, q) - b).max() < 1e-13 * n


def test_fplu_tie_goes_to_first_in_row_major_order():
    """Equal maxima at (0, 1) and (1, 0): the strict `>` row-by-row scan keeps
    (0, 1), so the first step swaps columns and not rows (a column-major scan
    would swap rows instead)."""
    a = np.array([[1.0, 5.0], [5.0, 2.0]])
    lu, p, q, prow, pcol, steps = orc.fplu_decompose(a)
    assert (prow[0], pcol[0]) == (0, 1)
    assert list(q) == [1, 0] and list(p) == [0, 1]
    # the diagonal (index, index) element is kept on a tie with it
    b = np.array([[-5.0, 5.0], [5.0, 1.0]])
    _, _, _, prow, pcol, _ = orc.fplu_decompose(b)
    assert (prow[0], pcol[0]) == (0, 0)


def test_fplu_zero_pivot_stops_elimination():
    """A zero trailing block ends the loop (lu.rs:405-407): nothing after it is
    swapped or eliminated, and rank counts the steps taken."""
    a = np.zeros((4, 4))
    a[:2, :2] = [[2.0, 1.0], [1.0, 3.0]]
    a[3, 0] = 0.0
    lu, p, q, prow, pcol, steps = orc.fplu_decompose(a)
    assert steps == 2 and orc.fplu_rank(lu) == 2 and not orc.fplu_is_invertible(lu)
    # rank 3 in an order-30 matrix with its nonzero block scattered by random
    # row and column permutations: the trailing block is exactly zero after 3 steps
    r = rng(14)
    b = np.zeros((30, 30))
    b[:3, :3] = r.standard_normal((3, 3))
    b = b[r.permutation(30)][:, r.permutation(30)]
    lu, p, q, prow, pcol, steps = orc.fplu_decompose(b)
    assert steps == 3 and orc.fplu_rank(lu) == 3


def test_fplu_det_and_solve_identities():
    """det = det(A).  solve follows lu.rs:458-464 literally: it maps the
    substitutions' result y through `&q * y` (x[q[i]] = y[i]) although
    P A Q = L U makes the solution x[q^-1[i]] = y[i].  The two agree when q is
    its own inverse (the reference's 4 x 4 tests); in general the returned x
    satisfies A x' = b for x'[j] = x[q[q[j]]].  A restatement that applied
    q^-1, or no column permutation, fails this test."""
    r = rng(15)
    for n in (1, 3, 12, 40):
        a = r.standard_normal((n, n))
        lu, p, q, *_ = orc.fplu_decompose(a)
        assert abs(orc.fplu_det(lu, p, q) - np.linalg.det(a)) <= 1e-12 * abs(np.linalg.det(a)) * n
        b = r.standard_normal(n)
        x = orc.fplu_solve(lu, p, q, b)
        xs = x[q[q]]
        assert np.abs(a @ xs - b).max() < 1e-12 * n * np.abs(b).max()
        if n >= 12:
            assert not np.array_equal(q[q], np.arange(n))  # the case above is not vacuous
        inv = orc.fplu_inverse(lu, p, q)
        assert np.abs(a @ inv[q[q]] - np.eye(n)).max() < 1e-11 * n
    # an involutive q: solve and inverse are the true ones
    a = np.array([[1.0, 5.0, 0.5], [0.25, 2.0, 1.0], [3.0, 0.5, 1.0]])
    lu, p, q, *_ = orc.fplu_decompose(a)
    assert np.array_equal(q[q], np.arange(3))
    assert np.abs(orc.fplu_inverse(lu, p, q) @ a - np.eye(3)).max() < 1e-15 * 8


def test_fplu_f32():
    a = rng(16).standard_normal((20, 20)).astype(np.float32)
    lu, p, q, prow, pcol, steps = orc.fplu_decompose(a, dtype=np.float32)
    assert lu.dtype == np.float32 and steps == 20
    lu64, _, _, prow64, pcol64, _ = orc.fplu_decompose(a.astype(np.float64))
    assert np.array_equal(prow, prow64) and np.array_equal(pcol, pcol64)
