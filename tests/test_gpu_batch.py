"""The batched engine (culip_amd.Batch) on the GPU: every system's solution,
residuals and iteration counts are bit for bit those of the same system
factored and solved alone by BlockiLDLt (block-Jacobi solves, same sweeps)
+ SQMR with the same options, whatever else is in the batch (the batched
refactor runs each level of all systems as one launch per step). Iteration limits that are
not a multiple of the true-residual period end runs between checks, and an
unattainable tolerance drives refinement restarts and undone corrections."""
import numpy as np
import pytest
import scipy.sparse as sp

import culip_amd as ca

pytestmark = pytest.mark.gpu


def _systems(sizes, B=8, seed0=300):
    out = []
    for i, n in enumerate(sizes):
        k = ca.KKT(n, int(0.4 * n), 4, 3, seed=seed0 + i)
        p = ca.Prep(k.row_ptr, k.col, k.val, n, B, ca.ORDER_BLOCK_COLOR)
        out.append((k, p, np.ascontiguousarray(k.rhs[p.perm])))
    return out


def _alone(p, vals, rhs, prec, sweeps, tol, max_it, period, restarts, dev):
    import torch
    g = ca.BlockiLDLt(p.row_ptr, p.col, vals, p.piv_starts, p.block_starts, prec)
    g.compute(0, 1.0, 1e-4)
    g.set_solve_algorithm("jacobi", sweeps)
    A = ca.CSRMatrixSPMV(p.row_ptr, p.col, vals)
    q = ca.SQMR(A, tol, max_it, g, true_residual_period=period, max_restarts=restarts)
    x = torch.zeros(rhs.size, dtype=torch.float64, device=dev)
    q.solve(torch.tensor(rhs, device=dev), x)
    return x.cpu().numpy(), dict(q.result)


def _check_equal(batch_out, alone_out, label):
    xb, rb = batch_out
    xa, ra = alone_out
    assert np.array_equal(xb, xa), (label, float(np.max(np.abs(xb - xa))))
    for key in ("iterations", "restarts", "converged", "true_rel_residual", "residual"):
        assert rb[key] == ra[key], (label, key, rb[key], ra[key])


def _run(dev, systems, prec, sweeps, tol=1e-8, max_it=1000, period=20, restarts=10,
         vals=None):
    bt = ca.Batch(prec=prec, tol=tol, max_it=max_it, true_residual_period=period,
                  max_restarts=restarts, jacobi_sweeps=sweeps)
    for _, p, _ in systems:
        bt.add(p.row_ptr, p.col, p.val, p.piv_starts, p.block_starts)
    vals = vals or [p.val for _, p, _ in systems]
    xs, res = bt.solve(vals, [r for _, _, r in systems])
    bt.close()
    out = []
    for i, ((_, p, rhs), v) in enumerate(zip(systems, vals)):
        a = _alone(p, v, rhs, prec, sweeps, tol, max_it, period, restarts, dev)
        _check_equal((xs[i], res[i]), a, (prec, sweeps, i))
        out.append(res[i])
    return out


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("sweeps", [2, 3])
def test_batch_equals_alone(hip_device, prec, sweeps):
    systems = _systems([1200, 2600, 4100, 1800, 3300, 5000])
    res = _run(hip_device, systems, prec, sweeps)
    for (k, p, rhs), r in zip(systems, res):
        assert r["true_rel_residual"] <= 1e-8, r


def test_batch_iteration_limit_between_checks(hip_device):
    # 30 iterations: one check at 20, then the run ends at its limit
    systems = _systems([1500, 3200, 2100])
    res = _run(hip_device, systems, "f64", 2, max_it=30)
    assert all(r["iterations"] == 30 for r in res), res


def test_batch_refinement_restarts(hip_device):
    # a tolerance no run can meet: runs end on stalls or drift, refinement
    # runs follow, and corrections that do not help are undone
    systems = _systems([1400, 2300, 3600, 2900], seed0=500)
    res = _run(hip_device, systems, "f32", 2, tol=1e-15, max_it=400, restarts=3)
    assert sum(r["restarts"] for r in res) > 0, res


def test_batch_new_values(hip_device):
    # two solves of one batch with different values: the second equals the
    # systems solved alone with those values
    systems = _systems([1700, 2800, 2200], seed0=700)
    bt = ca.Batch(prec="f32", jacobi_sweeps=2)
    for _, p, _ in systems:
        bt.add(p.row_ptr, p.col, p.val, p.piv_starts, p.block_starts)
    rhs = [r for _, _, r in systems]
    bt.solve([p.val for _, p, _ in systems], rhs)
    vals = [p.permute_values(k.barrier_values(1e-4, 1e4, seed=40 + i))
            for i, (k, p, _) in enumerate(systems)]
    xs, res = bt.solve(vals, rhs)
    info = bt.info()
    bt.close()
    assert info["systems"] == 3 and info["device_bytes"] > 0
    for i, ((k, p, b), v) in enumerate(zip(systems, vals)):
        a = _alone(p, v, b, "f32", 2, 1e-8, 1000, 20, 10, hip_device)
        _check_equal((xs[i], res[i]), a, i)
        K = sp.csr_matrix((v, p.col, p.row_ptr))
        assert np.linalg.norm(b - K @ xs[i]) / np.linalg.norm(b) <= 1e-8


def test_batch_rejects_bad_use(hip_device):
    with pytest.raises(ca.CulipError):
        ca.Batch(true_residual_period=0)
    (k, p, rhs), = _systems([1000])
    bt = ca.Batch(prec="f64")
    bt.add(p.row_ptr, p.col, p.val, p.piv_starts, p.block_starts)
    with pytest.raises(ValueError):
        bt.solve([], [])
    bt.solve([p.val], [rhs])
    with pytest.raises(ca.CulipError):
        bt.add(p.row_ptr, p.col, p.val, p.piv_starts, p.block_starts)
    bt.close()


def test_batch_device_arrays(hip_device):
    # values, right-hand sides and solutions in device memory: the same
    # results as with host arrays
    import torch
    systems = _systems([1300, 2400, 3100], seed0=900)
    bt = ca.Batch(prec="f64", jacobi_sweeps=3)
    for _, p, _ in systems:
        bt.add(p.row_ptr, p.col, p.val, p.piv_starts, p.block_starts)
    vals = [p.val for _, p, _ in systems]
    rhs = [r for _, _, r in systems]
    xh, rh = bt.solve(vals, rhs)
    dv = [torch.tensor(v, device=hip_device) for v in vals]
    db = [torch.tensor(r, device=hip_device) for r in rhs]
    xd, rd = bt.solve(dv, db)
    bt.close()
    for i in range(len(systems)):
        assert xd[i].is_cuda
        assert np.array_equal(xd[i].cpu().numpy(), xh[i])
        assert rd[i] == rh[i]


def test_batch_xcd_lists_equal_alone(hip_device):
    # 70 running systems: the product's and the sweeps' lists are XCD-aware
    # (each system's items on one XCD, holes between groups), same bits
    systems = _systems([600 + 37 * i for i in range(70)], seed0=1100)
    res = _run(hip_device, systems, "f32", 2)
    assert all(r["converged"] for r in res), res
