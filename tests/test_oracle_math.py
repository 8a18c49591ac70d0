"""Pins of the oracle's numeric factor and solves by what the mathematics
fixes: exact LDL' when nothing is dropped, a textbook dense LDL' with the
same 1x1/2x2 pivots, diagonal matrices, the tiny-pivot rules, the dropping
threshold, and inverse identities of the solves."""
import numpy as np
import pytest

import oracle_lib as ol
from conftest import sym_problem


def _dense(csr, n):
    rp, ci, cv = csr
    A = np.zeros((n, n))
    for i in range(n):
        A[i, ci[rp[i]:rp[i + 1]]] = cv[rp[i]:rp[i + 1]]
    return A


def _factor_dense(o, n):
    """U (unit upper, pivot blocks' upper parts) and D from the oracle."""
    bs = list(o.get("block_starts"))
    fb = o.get("fblocks").reshape(-1, 9)
    fix, fv = o.get("fix_store"), o.get("fval_store")
    U = np.zeros((n, n))
    for blk in fb:
        _, fmt, nnz, _, _, br, bc, cp, vp = blk
        r0, r1, c0, c1 = bs[br], bs[br + 1], bs[bc], bs[bc + 1]
        h, w = r1 - r0, c1 - c0
        if fmt == 1:
            T = fv[vp:vp + h * w].reshape(h, w)
            U[r0:r1, c0:c1] = np.triu(T) if br == bc else T
        else:
            for q in range(nnz):
                U[r0 + fix[cp + q] // w, c0 + fix[cp + q] % w] = fv[vp + q]
    tri = o.get("tri")
    D = np.zeros((n, n))
    for i in range(n):
        D[i, i] = tri[3 * i + 1]
        if i + 1 < n:
            D[i + 1, i] = tri[3 * i + 2]
            D[i, i + 1] = tri[3 * (i + 1)]
    return U, D


def _textbook_ldlt(A, is_piv):
    """Right-looking U'DU with the given 1x1/2x2 pivot starts (numpy)."""
    n = A.shape[0]
    S = np.triu(A) + np.triu(A, 1).T
    U = np.eye(n)
    D = np.zeros((n, n))
    p = 0
    while p < n:
        k = 2 if p + 1 < n and not is_piv[p + 1] else 1
        P = S[p:p + k, p:p + k]
        D[p:p + k, p:p + k] = P
        R = np.linalg.solve(P, S[p:p + k, p + k:])
        U[p:p + k, p + k:] = R
        S[p + k:, p + k:] -= R.T @ P @ R
        p += k
    return U, D


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_single_block_equals_textbook_ldlt(seed):
    csr, piv, bs, A = sym_problem(n_blocks=1, bmax=32, density=0.5, seed=seed)
    n = A.shape[0]
    o = ol.OracleBildlt(*csr, piv, [0], "d")
    o.symbolic(0, 1.0)
    o.factor(0.0)
    U, D = _factor_dense(o, n)
    is_piv = np.zeros(n, bool)
    is_piv[piv] = True
    Ut, Dt = _textbook_ldlt(A, is_piv)
    np.testing.assert_allclose(D, Dt, rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(np.triu(U, 1), np.triu(Ut, 1), rtol=1e-8, atol=1e-8)
    np.testing.assert_allclose(U.T @ D @ U, A, rtol=1e-9, atol=1e-9)


@pytest.mark.parametrize("t", ["d", "f"])
def test_complete_block_fill_is_exact(t):
    """Dense pattern: every block is dense, nothing is dropped, U'DU = A."""
    csr, piv, bs, A = sym_problem(n_blocks=5, bmax=12, density=1.0, seed=11)
    n = A.shape[0]
    o = ol.OracleBildlt(*csr, piv, bs, t)
    o.symbolic(0, 1.0)
    o.factor(1e-4)
    fb = o.get("fblocks").reshape(-1, 9)
    assert (fb[:, 1] == 1).all()
    U, D = _factor_dense(o, n)
    tol = 1e-9 if t == "d" else 2e-3
    np.testing.assert_allclose(U.T @ D @ U, A, rtol=tol, atol=tol * np.abs(A).max())


def test_diagonal_matrix():
    n = 40
    d = np.linspace(-3, 5, n)
    d[np.abs(d) < 0.1] = 0.7
    csr = (np.arange(n + 1, dtype=np.int32), np.arange(n, dtype=np.int32), d)
    o = ol.OracleBildlt(*csr, np.arange(n), [0, 16, 32], "d")
    o.symbolic(0, 1.0)
    o.factor(1e-4)
    U, D = _factor_dense(o, n)
    np.testing.assert_array_equal(np.diag(D), d)
    np.testing.assert_array_equal(U, np.eye(n))


def test_tiny_pivot_rules():
    """1x1 |d| < 1e-8 -> d = 1e-6, U(p,p) = 1 (kernels :556-580); 2x2
    |det| < 1e-8 -> d11 = d22 = 1e-6, d12 kept, U diagonal block = I
    (:609-663)."""
    A = np.diag([1e-9, 2.0, 1e-5, 1e-5, 3.0])
    A[2, 3] = A[3, 2] = 1e-9   # det = 1e-10 - 1e-18 < 1e-8
    A[0, 4] = A[4, 0] = 0.5
    n = 5
    rp, ci, cv = [0], [], []
    for i in range(n):
        nz = np.nonzero(A[i])[0]
        ci += list(nz)
        cv += list(A[i, nz])
        rp.append(len(ci))
    piv = [0, 1, 2, 4]   # 2x2 at rows 2, 3
    o = ol.OracleBildlt(np.array(rp), np.array(ci), np.array(cv), piv, [0], "d")
    o.symbolic(0, 1.0)
    o.factor(0.0)
    tri = o.get("tri")
    assert tri[1] == 1e-6                      # replaced 1x1 pivot, sign lost
    assert tri[3 * 2 + 1] == 1e-6 and tri[3 * 3 + 1] == 1e-6
    assert tri[3 * 2 + 2] == 1e-9 and tri[3 * 3] == 1e-9
    U, D = _factor_dense(o, n)
    assert U[0, 0] == 1.0 and U[2, 2] == 1.0 and U[3, 3] == 1.0 and U[2, 3] == 0.0
    assert U[0, 4] == 0.5 / 1e-6               # row scaled by the replaced pivot


def test_threshold_dropping_rule():
    """An off-diagonal sparse block entry is dropped when |u| < tau *
    sqrtf(sum of squares of its row of U) (kernels :1919-1946)."""
    n, nb = 64, 2
    A = np.eye(n) * 2.0
    # row 0 couples to columns 40 (large) and 41 (small)
    A[0, 40] = A[40, 0] = 1.0
    A[0, 41] = A[41, 0] = 1e-5
    csr = _to_csr(A)
    for tau, kept in ((1e-6, 2), (1e-3, 1)):
        o = ol.OracleBildlt(*csr, np.arange(n), [0, 32], "d")
        o.symbolic(0, 4.0)
        o.factor(tau)
        fb = o.get("fblocks").reshape(-1, 9)
        off = fb[(fb[:, 5] == 0) & (fb[:, 6] == 1)][0]
        assert off[1] == 0 and off[2] == kept
        rn = o.get("row_norms")
        assert np.isclose(rn[0], 1.0 + 0.25 + (2.5e-11 if kept == 2 else 2.5e-11))


def _to_csr(A):
    n = A.shape[0]
    rp, ci, cv = [0], [], []
    for i in range(n):
        nz = np.nonzero(A[i])[0]
        ci += list(nz)
        cv += list(A[i, nz])
        rp.append(len(ci))
    return np.array(rp, np.int32), np.array(ci, np.int32), np.array(cv)


@pytest.mark.parametrize("t", ["d", "f"])
def test_solve_inverse_identities(t):
    csr, piv, bs, A = sym_problem(n_blocks=8, bmax=20, density=0.1, seed=5)
    n = A.shape[0]
    o = ol.OracleBildlt(*csr, piv, bs, t)
    o.symbolic(1, 2.0)
    o.factor(1e-4)
    U, D = _factor_dense(o, n)
    b = np.random.default_rng(2).normal(size=n)
    tol = 1e-10 if t == "d" else 1e-3
    xr = o.solve("right", b)
    assert np.linalg.norm(U @ xr - b) / np.linalg.norm(b) < tol * np.linalg.cond(U)
    xl = o.solve("left", b)
    assert np.linalg.norm(U.T @ xl - b) / np.linalg.norm(b) < tol * np.linalg.cond(U)
    xm = o.solve("middle", b)
    assert np.linalg.norm(D @ xm - b) / np.linalg.norm(b) < tol * 10
