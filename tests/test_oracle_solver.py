"""Pins of the oracle's solver arithmetic (CPU only): its dot product (BLAS::dot, BLAS.cu:98-104)
against the exactly rounded sum and known answers.  tests/test_oracle_mutants.py plants mistakes in
it and checks that this file fails."""
import numpy as np

from oracle import oracle as O


def test_solver_dot_product_exact_sum():
    # BLAS::dot (BLAS.cu:98-104) of the oracle's solver.  A float x float product is exact in double, so
    # the FP64-accumulated sum differs from the exactly rounded sum (math.fsum) by at most n ulps of the
    # sum of |a_i c_i|; the FP32 instance rounds that once to float.  Known answers: orthogonal vectors
    # give 0, and sum_{i<n} i * 1 = n (n - 1) / 2 exactly.
    import math
    rng = np.random.default_rng(3)
    for n in (1, 7, 1000, 100003):
        a = rng.standard_normal(n).astype(np.float32)
        c = rng.standard_normal(n).astype(np.float32)
        prods = [float(x) * float(y) for x, y in zip(a, c)]
        exact = math.fsum(prods)
        scale = math.fsum(abs(x) for x in prods)
        got64 = O.dot(a.astype(np.float64), c.astype(np.float64))
        assert abs(got64 - exact) <= n * 2.3e-16 * scale, (n, got64, exact)
        got32 = O.dot(a, c)
        assert abs(got32 - np.float32(exact)) <= 2 * np.spacing(np.float32(abs(exact))) + n * 2.3e-16 * scale, n
    n = 4096
    i = np.arange(n, dtype=np.float32)
    assert O.dot(i, np.ones(n, np.float32)) == n * (n - 1) / 2
    e0 = np.zeros(n, np.float64)
    e1 = np.zeros(n, np.float64)
    e0[::2] = 1.0
    e1[1::2] = 1.0
    assert O.dot(e0, e1) == 0.0
    assert O.dot(e0, e0) == n / 2
