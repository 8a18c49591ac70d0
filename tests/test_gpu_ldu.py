"""GPU: ImportLDUPreconditioner and SQMR on the reference's own fixtures.

The solves are checked against dense textbook triangular solves (scipy) of
the same factors, with the permutation / scaling conventions of solve_PS
(import_ldu_preconditioner.impl.cuh:414-484: pull x = P'(s .* b), push
x = s .* (P b)). SQMR with the imported factor runs the sqmr_{small,big}
fixtures to the reference test's bounds (test_sqmr.cc:262-267:
||b - Ax|| <= 1e-5 ||b|| for double and 1e-4 for float, iterations <= n)
and is compared with the oracle's SQMR on the same inputs.
"""
import numpy as np
import pytest
import scipy.linalg as sla
import torch

import culip_amd as ca
import oracle_lib as ol
from conftest import ldu_problem

pytestmark = pytest.mark.gpu


def _csr(M):
    rp, ci, cv = [0], [], []
    for i in range(M.shape[0]):
        nz = np.nonzero(M[i])[0]
        ci.extend(nz)
        cv.extend(M[i, nz])
        rp.append(len(ci))
    return np.array(rp, np.int32), np.array(ci, np.int32), np.array(cv)


def _triangular(n, seed, lower=True, groups=None):
    """Random sparse triangular factor with a non-unit diagonal. With
    `groups`, rows fall into that many consecutive groups and only depend on
    earlier groups (at most `groups` dependency levels)."""
    rng = np.random.default_rng(seed)
    M = np.where(rng.random((n, n)) < 0.04, rng.normal(size=(n, n)) * 0.3, 0.0)
    if groups:
        g = np.arange(n) * groups // n
        M[g[:, None] <= g[None, :]] = 0.0
    M = np.tril(M, -1)
    M[np.arange(n), np.arange(n)] = rng.uniform(0.5, 2.0, n) * rng.choice([-1, 1], n)
    return M if lower else M.T.copy()


def _tri_d(n, seed):
    """Tridiagonal D layout {D(i-1,i), D(i,i), D(i+1,i)} with some 2x2 pivots
    and its dense matrix."""
    rng = np.random.default_rng(seed)
    tri = np.zeros(3 * n)
    is_piv = np.ones(n, np.int32)
    D = np.zeros((n, n))
    i = 0
    while i < n:
        if i + 1 < n and rng.random() < 0.3:
            a, c, b = rng.normal(size=3) + np.array([3.0, 0.0, -3.0])
            D[i, i], D[i + 1, i + 1], D[i, i + 1], D[i + 1, i] = a, b, c, c
            tri[3 * i + 1], tri[3 * i + 2] = a, c
            tri[3 * (i + 1)], tri[3 * (i + 1) + 1] = c, b
            is_piv[i + 1] = 0
            i += 2
        else:
            D[i, i] = rng.uniform(0.5, 2.0) * rng.choice([-1, 1])
            tri[3 * i + 1] = D[i, i]
            i += 1
    return tri, is_piv, D


def _pull(p, s, b):
    return s[p] * b[p]


def _push(p, s, b):
    x = np.empty_like(b)
    x[p] = s[p] * b
    return x


def _run(solve, b, dev):
    bd = torch.tensor(b, device=dev)
    xd = torch.empty_like(bd)
    solve(bd, xd)
    torch.cuda.synchronize()
    # in place as well (b == x)
    bd2 = torch.tensor(b, device=dev)
    solve(bd2, bd2)
    torch.cuda.synchronize()
    x1, x2 = xd.cpu().numpy(), bd2.cpu().numpy()
    np.testing.assert_array_equal(x1, x2)
    return x1


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_symmetric_solves_against_dense(hip_device, prec):
    n = 300
    L = _triangular(n, 1)
    tri, is_piv, D = _tri_d(n, 2)
    rng = np.random.default_rng(3)
    pr, pc = rng.permutation(n).astype(np.int32), rng.permutation(n).astype(np.int32)
    sr, sc = rng.uniform(0.5, 2.0, n), rng.uniform(0.5, 2.0, n)
    g = ca.ImportLDUPreconditioner(_csr(L), tri, is_piv, prec=prec)
    g.set_permutation(pr, pc)
    g.set_scaling(sr, sc)
    assert g.n() == n and g.is_middle()
    Lr = L.astype(np.float32).astype(np.float64) if prec == "f32" else L
    b = rng.normal(size=n)
    tol = dict(rtol=1e-11, atol=1e-11)
    # left: inv(L) P_r' S_r ; transposed: S_r P_r inv(L')
    ref = sla.solve_triangular(Lr, _pull(pr, sr, b), lower=True)
    np.testing.assert_allclose(_run(lambda u, v: g.solve_left(u, v), b, hip_device), ref, **tol)
    ref = _push(pr, sr, sla.solve_triangular(Lr, b, lower=True, trans="T"))
    np.testing.assert_allclose(_run(lambda u, v: g.solve_left(u, v, True), b, hip_device), ref,
                               **tol)
    # right: S_c P_c inv(L') ; transposed: inv(L) P_c' S_c
    ref = _push(pc, sc, sla.solve_triangular(Lr, b, lower=True, trans="T"))
    np.testing.assert_allclose(_run(lambda u, v: g.solve_right(u, v), b, hip_device), ref, **tol)
    ref = sla.solve_triangular(Lr, _pull(pc, sc, b), lower=True)
    np.testing.assert_allclose(_run(lambda u, v: g.solve_right(u, v, True), b, hip_device), ref,
                               **tol)
    # middle: the 1x1 / 2x2 block-diagonal D
    np.testing.assert_allclose(_run(lambda u, v: g.solve_middle(u, v), b, hip_device),
                               np.linalg.solve(D, b), **tol)
    # with both removed the outer steps are the identity
    g.set_permutation(None)
    g.set_scaling(None)
    ref = sla.solve_triangular(Lr, b, lower=True)
    np.testing.assert_allclose(_run(lambda u, v: g.solve_left(u, v), b, hip_device), ref, **tol)


def test_lu_and_cholesky_forms(hip_device):
    n = 250
    L, U = _triangular(n, 4), _triangular(n, 5, lower=False)
    rng = np.random.default_rng(6)
    b = rng.normal(size=n)
    tol = dict(rtol=1e-11, atol=1e-11)
    g = ca.ImportLDUPreconditioner(_csr(L), U=_csr(U))
    assert not g.is_middle()
    np.testing.assert_allclose(_run(lambda u, v: g.solve_right(u, v), b, hip_device),
                               sla.solve_triangular(U, b, lower=False), **tol)
    np.testing.assert_allclose(_run(lambda u, v: g.solve_right(u, v, True), b, hip_device),
                               sla.solve_triangular(U, b, lower=False, trans="T"), **tol)
    np.testing.assert_allclose(_run(lambda u, v: g.solve_left(u, v), b, hip_device),
                               sla.solve_triangular(L, b, lower=True), **tol)
    # no D: the middle solve copies
    np.testing.assert_array_equal(_run(lambda u, v: g.solve_middle(u, v), b, hip_device), b)
    # the upper triangle of a full matrix given as L is ignored
    F = L + np.triu(rng.normal(size=(n, n)) * (rng.random((n, n)) < 0.05), 1)
    h = ca.ImportLDUPreconditioner(_csr(F))
    np.testing.assert_allclose(_run(lambda u, v: h.solve_left(u, v), b, hip_device),
                               sla.solve_triangular(L, b, lower=True), **tol)
    np.testing.assert_allclose(_run(lambda u, v: h.solve_right(u, v), b, hip_device),
                               sla.solve_triangular(L, b, lower=True, trans="T"), **tol)


def test_approximate_solves_exact_on_shallow_factors(hip_device):
    """exact_solve=False: x0 = D^-1 b, then 10 Jacobi sweeps. On a triangular
    factor with at most 8 dependency levels the iteration matrix is
    nilpotent of that index, so the result equals the exact solve."""
    n = 400
    L = _triangular(n, 7, groups=8)
    rng = np.random.default_rng(8)
    b = rng.normal(size=n)
    g = ca.ImportLDUPreconditioner(_csr(L), exact_solve=False)
    tol = dict(rtol=1e-10, atol=1e-10)
    np.testing.assert_allclose(_run(lambda u, v: g.solve_left(u, v), b, hip_device),
                               sla.solve_triangular(L, b, lower=True), **tol)
    np.testing.assert_allclose(_run(lambda u, v: g.solve_right(u, v), b, hip_device),
                               sla.solve_triangular(L, b, lower=True, trans="T"), **tol)
    # a deep factor is only approximated: the result differs from the exact one
    Ld = _triangular(n, 9)
    Ld[np.arange(1, n), np.arange(n - 1)] = 0.9
    h = ca.ImportLDUPreconditioner(_csr(Ld), exact_solve=False)
    xa = _run(lambda u, v: h.solve_left(u, v), b, hip_device)
    assert np.abs(xa - sla.solve_triangular(Ld, b, lower=True)).max() > 1e-6


@pytest.mark.parametrize("prefix,precondition,prec", [
    ("sqmr_small", False, "f64"), ("sqmr_small", True, "f64"), ("sqmr_small", True, "f32"),
    ("sqmr_big", True, "f64"), ("sqmr_big", True, "f32")])
def test_sqmr_reference_fixtures(hip_device, golden, prefix, precondition, prec):
    (ar, ac, av), Lc, tri, is_piv, p, s, n = ldu_problem(golden, prefix)
    A = ca.CSRMatrixSPMV(ar, ac, av)
    b = ol.csr_spmv(ar, ac, av, np.ones(n))
    g = None
    if precondition:
        g = ca.ImportLDUPreconditioner(Lc, tri, is_piv, prec=prec)
        g.set_permutation(p, p)
        g.set_scaling(s, s)
    # test_sqmr.cc:281-292: term_tol 1e-10 / test_tol 1e-5 (double),
    # 1e-6 / 1e-4 (float); max_it = n
    term, test = (1e-10, 1e-5) if prec == "f64" else (1e-6, 1e-4)
    solver = ca.SQMR(A, term, n, g)
    bd = torch.tensor(b, device=hip_device)
    xd = torch.zeros_like(bd)
    solver.solve(bd, xd)
    x = xd.cpu().numpy()
    r = solver.result
    rel = np.linalg.norm(b - ol.csr_spmv(ar, ac, av, x)) / np.linalg.norm(b)
    print(prefix, precondition, prec, "iterations", r["iterations"], "rel", rel)
    assert rel <= test, rel
    assert r["iterations"] <= n
    if prec == "f64":
        xo, ro = ol.sqmr_ldu(ar, ac, av, Lc if precondition else None, tri, is_piv, p, s, b,
                             term, n)
        # same algorithm, different summation order: iteration counts agree
        # up to the four-iteration test throttle
        assert abs(r["iterations"] - ro.iterations) <= 4, (r["iterations"], ro.iterations)
        np.testing.assert_allclose(x, xo, rtol=0, atol=1e-6 * np.abs(xo).max())
