"""The oracle against the reference's own fixtures (tests/la/data, copied to
tests/golden) and their known answers.

spmv_A/d/K pin the KKT operator (test_spmv.cc:191-210: K = [diag(d) A'; A 0]
exactly). sqmr_{small,big}_* pin SQMR with the ImportLDU preconditioner
(test_sqmr.cc:192-293: ||b - Ax|| <= 1e-5 ||b||, iterations <= n, double).
"""
import numpy as np
import pytest
import scipy.sparse as sp

import culip_amd as ca
import oracle_lib as ol
from conftest import ldu_problem


def _csr(path):
    rp, ci, cv, n = ca.read_matrix_csr(path)
    return rp, ci, cv, n


def test_kkt_fixture_is_the_augmented_operator(golden):
    ar, ac, av, an = _csr(f"{golden}/spmv_A.mtx")
    d = ca.read_dense_vector(f"{golden}/spmv_d.mtx")
    kr, kc, kv, kn = _csr(f"{golden}/spmv_K.mtx")
    m = ar.size - 1
    assert kr.size - 1 == an + m
    rng = np.random.default_rng(0)
    for x in (np.full(an + m, 0.1), rng.normal(size=an + m)):
        yk = ol.csr_spmv(kr, kc, kv, x)
        ya = ol.kkt_spmv(ar, ac, av, an, d, 0.0, None, x)
        np.testing.assert_allclose(ya, yk, rtol=1e-10, atol=1e-10)
    # scipy as an independent textbook SpMV
    K = sp.csr_matrix((kv, kc, kr), shape=(an + m, an + m))
    x = rng.normal(size=an + m)
    np.testing.assert_allclose(ol.csr_spmv(kr, kc, kv, x), K @ x, rtol=1e-12, atol=1e-12)


def test_kkt_adapter_alpha_beta_delta_ix():
    rng = np.random.default_rng(1)
    m, n = 7, 11
    A = sp.random(m, n, density=0.4, random_state=2, format="csr")
    d = rng.uniform(1, 2, size=n)
    ix = np.array([1, 4, 5])
    delta = -3e-2
    K = sp.bmat([[sp.diags(d), A.T], [A, None]]).toarray()
    for i in ix:
        K[n + i, n + i] += delta
    x, y0 = rng.normal(size=n + m), rng.normal(size=n + m)
    out = ol.kkt_spmv(A.indptr, A.indices, A.data, n, d, delta, ix, x, y0, 2.0, -0.5)
    np.testing.assert_allclose(out, 2.0 * K @ x - 0.5 * y0, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("prefix,precondition,two_by_two", [
    ("sqmr_small", False, 0), ("sqmr_small", True, 0), ("sqmr_big", True, 15)])
def test_sqmr_reference_fixtures(golden, prefix, precondition, two_by_two):
    (ar, ac, av), L, tri, is_piv, p, s, n = ldu_problem(golden, prefix)
    assert (is_piv == 0).sum() == two_by_two
    x_true = np.ones(n)
    A = sp.csr_matrix((av, ac, ar), shape=(n, n))
    b = A @ x_true
    x, r = ol.sqmr_ldu(ar, ac, av, L if precondition else None, tri, is_piv, p, s, b,
                       1e-10, n)
    rel = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
    assert rel <= 1e-5, rel                  # test_sqmr.cc:266
    assert r.iterations <= n                 # test_sqmr.cc:267
    if precondition:
        # an exact LDL' preconditioner converges in a handful of iterations
        assert r.iterations <= 12, r.iterations


def test_sqmr_identity_preconditioner_finite_termination():
    """Without preconditioning SQMR on an n x n symmetric system reaches a
    tiny residual in about n steps (exact Krylov termination)."""
    rng = np.random.default_rng(3)
    n = 24
    Q = rng.normal(size=(n, n))
    A = Q + Q.T + np.diag(np.linspace(-5, 5, n))
    S = sp.csr_matrix(A)
    b = rng.normal(size=n)
    x, r = ol.sqmr_ldu(S.indptr, S.indices, S.data, None, None, None, None, None, b, 1e-12,
                       3 * n)
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) < 1e-8
    assert r.iterations <= n + 8


def test_solve_D_against_dense_blocks():
    """k_solve_D restatement equals a dense solve with the 1x1/2x2 blocks."""
    rng = np.random.default_rng(4)
    n = 9
    is_piv = np.array([1, 1, 0, 1, 1, 0, 1, 0, 1], np.int32)
    tri = np.zeros(3 * n)
    D = np.zeros((n, n))
    i = 0
    while i < n:
        if i + 1 < n and not is_piv[i + 1]:
            a, c, e = rng.normal(size=3) * 3
            tri[3 * i + 1], tri[3 * i + 2] = a, c
            tri[3 * (i + 1)], tri[3 * (i + 1) + 1] = c, e
            D[i:i + 2, i:i + 2] = [[a, c], [c, e]]
            i += 2
        else:
            tri[3 * i + 1] = rng.normal() * 3
            D[i, i] = tri[3 * i + 1]
            i += 1
    b = rng.normal(size=n)
    np.testing.assert_allclose(ol.solve_D(tri, is_piv, b), np.linalg.solve(D, b), rtol=1e-12)


def test_normal_fixture_is_a_dinv_at(golden):
    """spmv_N.mtx is A diag(d)^-1 A' for spmv_A / spmv_d (the operator of
    NormalMatrixSPMV, spmv.impl.cuh:214-240): scipy, to the fixture's digits."""
    ar, ac, av, an = _csr(f"{golden}/spmv_A.mtx")
    d = ca.read_dense_vector(f"{golden}/spmv_d.mtx")
    nr, nc, nv, nn = _csr(f"{golden}/spmv_N.mtx")
    A = sp.csr_matrix((av, ac, ar), shape=(ar.size - 1, an))
    N = sp.csr_matrix((nv, nc, nr), shape=(nr.size - 1, nn))
    M = A @ sp.diags(1.0 / d) @ A.T
    assert abs(M - N).max() <= 1e-12 * abs(N).max()
    # and not A diag(d) A'
    assert abs(A @ sp.diags(d) @ A.T - N).max() > 1.0
