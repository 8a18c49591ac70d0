"""Symmetric equilibration on the device and the solve of the scaled
system against the unscaled residual (the serving path's input
preparation for barrier-conditioned KKT systems).

The device scaling is checked bit for bit against a numpy restatement of
the same sweeps (maxima are exact; sqrt and division correctly rounded in
both); the scaled solve equals, bit for bit, the one-system path composed
by hand from the restated scaling; and every request of a barrier-weighted
serving stream meets 1e-8 on the unscaled system."""
import numpy as np
import pytest
import scipy.sparse as sp

import culip_amd as ca
import culip_amd.batch as cb

pytestmark = pytest.mark.gpu

SWEEPS = 8


def ruiz(rp, ci, v, sweeps):
    """Restatement: per sweep, for all rows at once, m_i = max_j
    |(d_i a_ij) d_j| and d_i <- d_i / sqrt(m_i) (kept where m_i = 0);
    returns (d, scaled values (d_i a_ij) d_j)."""
    n = rp.size - 1
    rows = np.repeat(np.arange(n), np.diff(rp))
    d = np.ones(n)
    for _ in range(sweeps):
        prod = np.abs((d[rows] * v) * d[ci])
        m = np.zeros(n)
        np.maximum.at(m, rows, prod)
        with np.errstate(divide="ignore"):
            d = np.where(m > 0, d / np.sqrt(m), d)
    return d, (d[rows] * v) * d[ci]


def _barrier_system(n, seed, wseed):
    k, p = cb.kkt_system(n, seed, 32)
    vals = p.permute_values(k.barrier_values(1e-4, 1e4, seed=wseed))
    return k, p, vals


def test_device_equilibration_equals_restatement(hip_device):
    import torch
    k, p, vals = _barrier_system(20000, 7, 3)
    op = ca.CSRMatrixSPMV(p.row_ptr, p.col, vals)
    N = p.row_ptr.size - 1
    d = torch.zeros(N, dtype=torch.float64, device=hip_device)
    op.equilibrate(SWEEPS, d)
    torch.cuda.synchronize()
    dr, sv = ruiz(p.row_ptr, p.col, vals, SWEEPS)
    assert np.array_equal(d.cpu().numpy(), dr)
    # the operator now applies D K D
    x = np.random.default_rng(1).normal(size=N)
    y = torch.zeros(N, dtype=torch.float64, device=hip_device)
    op.multiply(torch.tensor(x, device=hip_device), y)
    ref = sp.csr_matrix((sv, p.col, p.row_ptr)) @ x
    np.testing.assert_allclose(y.cpu().numpy(), ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
    S = sp.csr_matrix((sv, p.col, p.row_ptr))
    assert abs(S).max() <= 1.0 + 1e-12


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_served_scaled_solve_equals_composed_path(hip_device, prec):
    """Server with equilibration == restated scaling + BlockiLDLt + SQMR on
    D K D y = D b with the residual measured on K, then x = D y."""
    import torch
    k, p, vals = _barrier_system(30000, 8, 4)
    rhs = k.rhs[p.perm].copy()
    s = ca.Server(slots=2, prec=prec, max_it=3000, solve_algorithm="jacobi", jacobi_sweeps=2,
                  equilibrate_sweeps=SWEEPS)
    pid = s.add_pattern(p.row_ptr, p.col, p.val, p.piv_starts, p.block_starts)
    x_srv, res, _, _ = s.wait(s.submit(pid, vals, rhs))
    s.close()
    d, sv = ruiz(p.row_ptr, p.col, vals, SWEEPS)
    g = ca.BlockiLDLt(p.row_ptr, p.col, sv, p.piv_starts, p.block_starts, prec)
    g.compute(0, 1.0, 1e-4)
    g.set_solve_algorithm("jacobi", 2)
    A = ca.CSRMatrixSPMV(p.row_ptr, p.col, sv)
    q = ca.SQMR(A, 1e-8, 3000, g, true_residual_period=20, max_restarts=10)
    dd = torch.tensor(d, device=hip_device)
    q.set_residual_scale(dd)
    y = torch.zeros(rhs.size, dtype=torch.float64, device=hip_device)
    q.solve(torch.tensor(d * rhs, device=hip_device), y)
    x = d * y.cpu().numpy()
    assert np.array_equal(x, x_srv)
    assert q.result == res
    K = sp.csr_matrix((vals, p.col, p.row_ptr))
    rel = np.linalg.norm(rhs - K @ x) / np.linalg.norm(rhs)
    assert rel <= 1e-8 and abs(rel - res["true_rel_residual"]) <= 1e-6 * rel + 1e-12


def test_barrier_stream_converges_with_equilibration(hip_device):
    """Requests over three patterns (n = 20,000 to 100,000) with fresh
    barrier weights in [1e-4, 1e4]: every one meets 1e-8 on K."""
    pats = [cb.kkt_system(n, 900 + j, 32) for j, n in enumerate((20000, 60000, 100000))]
    s = ca.Server(slots=4, prec="f32", max_it=3000, solve_algorithm="jacobi", jacobi_sweeps=2,
                  equilibrate_sweeps=SWEEPS)
    pids = [s.add_pattern(p.row_ptr, p.col, p.val, p.piv_starts, p.block_starts)
            for _, p in pats]
    reqs = []
    for r in range(6):
        j = r % 3
        k, p = pats[j]
        vals = p.permute_values(k.barrier_values(1e-4, 1e4, seed=300 + r))
        reqs.append((j, vals, k.rhs[p.perm].copy()))
    ids = [s.submit(pids[j], v, b) for j, v, b in reqs]
    outs = [s.wait(i) for i in ids]
    s.close()
    for (j, vals, rhs), (x, res, _, _) in zip(reqs, outs):
        p = pats[j][1]
        K = sp.csr_matrix((vals, p.col, p.row_ptr))
        rel = np.linalg.norm(rhs - K @ x) / np.linalg.norm(rhs)
        assert rel <= 1e-8, (j, rel, res)
