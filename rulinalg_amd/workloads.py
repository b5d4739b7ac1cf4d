"""Seeded synthetic inputs of the benchmark workloads.

One generator for everything: numpy's Philox4x32-10 bit generator
(Salmon et al., SC'11) seeded with a small integer, so the CPU oracle and the
GPU see bit-identical inputs.  uniform(-1, 1) and standard normal draws use
numpy's Generator.uniform / Generator.standard_normal on that bit generator.

  order 256:          GEMM A ~ U[-1,1] seed 0, B ~ U[-1,1] seed 1;
                       LU A ~ N(0,1) seed 2, b ~ N(0,1) seed 3
  order 8192 / 16384: DGEMM 8192^3 U[-1,1] seeds 10, 11; LU 16384 N(0,1)
                       seed 12 with 16 RHS seed 13; Cholesky 16384
                       A = G G^T / n + I, G ~ N(0,1) seed 14
  mixed, order 16384: A = U diag(s) V^T with s log-spaced on [1, kappa],
                       U, V orthogonal (seeds 20, 21), x_true seed 22
"""
import numpy as np

ORDER256 = {"n": 256, "gemm_a": 0, "gemm_b": 1, "lu_a": 2, "lu_b": 3}
ORDER16384 = {"gemm_n": 8192, "gemm_a": 10, "gemm_b": 11, "lu_n": 16384, "lu_a": 12, "lu_b": 13,
              "lu_nrhs": 16, "chol_n": 16384, "chol_g": 14}
MIXED16384 = {"n": 16384, "u": 20, "v": 21, "x": 22, "kappa": 1e4, "kappa_fallback": 1e8}


def rng(seed):
    return np.random.Generator(np.random.Philox(seed))


def uniform(rows, cols, seed, dtype=np.float64):
    return rng(seed).uniform(-1.0, 1.0, (rows, cols)).astype(dtype, copy=False)


def gaussian(rows, cols, seed, dtype=np.float64):
    return rng(seed).standard_normal((rows, cols)).astype(dtype, copy=False)


def spd(n, seed):
    """A = G G^T / n + I with G ~ N(0,1) (host; small n only)."""
    g = gaussian(n, n, seed)
    return g @ g.T / n + np.eye(n)


def orthogonal(n, seed):
    """Q factor of a Gaussian matrix with the signs of R's diagonal folded in
    (a Haar-distributed orthogonal matrix).  Host; small n only."""
    q, r = np.linalg.qr(gaussian(n, n, seed))
    return q * np.sign(np.diag(r))


def conditioned(n, kappa, seed_u, seed_v):
    """A = U diag(s) V^T with s log-spaced from 1 to kappa (host; small n)."""
    s = np.logspace(0.0, np.log10(kappa), n)
    u = orthogonal(n, seed_u)
    v = orthogonal(n, seed_v)
    return (u * s) @ v.T
