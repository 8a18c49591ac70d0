"""The KKT operator with a general (CSR) Hessian block, matrix-free:
K = [H A'; A delta I(ix)] from H and A (culip_spmv_create_kkt_csr). With
H = diag(d) it is the reference's AugmentedMatrixSPMV, pinned by
tests/la/data/spmv_K.mtx (test_spmv.cc:191-210); with the KKT generator's
H it equals the assembled K; SQMR over it solves the system as SQMR over
the assembled matrix does."""
import numpy as np
import pytest
import scipy.sparse as sp

import culip_amd as ca
import oracle_lib as ol

pytestmark = pytest.mark.gpu


def test_diagonal_h_matches_fixture(hip_device, golden):
    import torch
    ar, ac, av, an = ca.read_matrix_csr(f"{golden}/spmv_A.mtx")
    d = ca.read_dense_vector(f"{golden}/spmv_d.mtx")
    kr, kc, kv, kn = ca.read_matrix_csr(f"{golden}/spmv_K.mtx")
    N = kr.size - 1
    n = an
    H = sp.csr_matrix((d, (np.arange(n), np.arange(n))), shape=(n, n))
    op = ca.AugmentedMatrixSPMV(ar, ac, av, an, H=(H.indptr, H.indices, H.data))
    assert op.m() == op.n() == N
    x = torch.full((N,), 0.1, dtype=torch.float64, device=hip_device)
    y = torch.zeros_like(x)
    op.multiply(x, y)
    # the diagonal form adds d_i x_i after the sum of (A' x2)_i, the CSR
    # form starts from H's entry: equal to rounding
    yd = torch.zeros_like(x)
    ca.AugmentedMatrixSPMV(ar, ac, av, an, d).multiply(x, yd)
    torch.cuda.synchronize()
    ref = ol.csr_spmv(kr, kc, kv, np.full(N, 0.1))
    np.testing.assert_allclose(y.cpu().numpy(), ref, rtol=1e-10, atol=1e-10)
    np.testing.assert_allclose(y.cpu().numpy(), yd.cpu().numpy(), rtol=1e-13, atol=1e-13)


def _blocks(n, m, seed=0):
    k = ca.KKT(n, m, 4, 3, seed=seed)
    K = sp.csr_matrix((k.val, k.col, k.row_ptr))
    H = sp.csr_matrix(K[:n, :n])
    A = sp.csr_matrix(K[n:, :n])
    for M in (H, A):
        M.sort_indices()
    delta = -float(K[n, n])
    return k, K, H, A, delta


def test_general_h_equals_assembled(hip_device):
    import torch
    n, m = 4000, 1600
    k, K, H, A, delta = _blocks(n, m)
    op = ca.AugmentedMatrixSPMV(A.indptr, A.indices, A.data, n, delta=-delta,
                                ix=np.arange(m), H=(H.indptr, H.indices, H.data))
    rng = np.random.default_rng(2)
    x = rng.normal(size=n + m)
    y0 = rng.normal(size=n + m)
    y = torch.tensor(y0, device=hip_device)
    op.multiply(torch.tensor(x, device=hip_device), y, alpha=1.5, beta=-0.25)
    torch.cuda.synchronize()
    ref = 1.5 * (K @ x) - 0.25 * y0
    np.testing.assert_allclose(y.cpu().numpy(), ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


def test_sqmr_over_matrix_free_kkt(hip_device):
    """BILDLt from the assembled K, SQMR with the matrix-free operator: the
    true residual of K meets 1e-8 and x agrees with SQMR over the assembled
    CSR operator (different summation orders: 1e-7 relative)."""
    import torch
    n, m = 2000, 800
    k, K, H, A, delta = _blocks(n, m)
    p = ca.Prep(k.row_ptr, k.col, k.val, n, 16, ca.ORDER_NATURAL)
    assert np.array_equal(p.perm, np.arange(n + m)), "natural ordering keeps K's layout"
    g = ca.BlockiLDLt(p.row_ptr, p.col, p.val, p.piv_starts, p.block_starts, "f64")
    g.compute(0, 1.0, 1e-4)
    op = ca.AugmentedMatrixSPMV(A.indptr, A.indices, A.data, n, delta=-delta,
                                ix=np.arange(m), H=(H.indptr, H.indices, H.data))
    b = k.rhs
    db = torch.tensor(b, device=hip_device)
    out = []
    for A_op in (op, ca.CSRMatrixSPMV(k.row_ptr, k.col, k.val)):
        q = ca.SQMR(A_op, 1e-8, 1000, g, true_residual_period=20, max_restarts=10)
        x = torch.zeros_like(db)
        q.solve(db, x)
        out.append((x.cpu().numpy(), dict(q.result)))
    (xm, rm), (xa, ra) = out
    rel = np.linalg.norm(b - K @ xm) / np.linalg.norm(b)
    assert rel <= 1e-8 and rm["true_rel_residual"] <= 1e-8, (rel, rm)
    assert abs(rm["iterations"] - ra["iterations"]) <= 20
    assert np.linalg.norm(xm - xa) / np.linalg.norm(xa) < 1e-7
