"""Request loop on the GPU: every result equals the result of the same
request solved alone (bit for bit: the same kernels run in the same order on
private buffers, whatever else is in flight), equals the single-system
BlockiLDLt + SQMR path, meets the FP64 true-residual tolerance, and the
device-memory budget is enforced."""
import numpy as np
import pytest
import scipy.sparse as sp

import culip_amd as ca

pytestmark = pytest.mark.gpu


def _patterns(sizes, B=16):
    out = []
    for i, n in enumerate(sizes):
        k = ca.KKT(n, int(0.4 * n), 4, 3, seed=100 + i)
        p = ca.Prep(k.row_ptr, k.col, k.val, n, B, ca.ORDER_BLOCK_COLOR)
        out.append((k, p))
    return out


def _request(k, p, seed):
    vals = p.permute_values(k.barrier_values(1e-4, 1e4, seed=seed))
    rhs = np.random.default_rng(seed).normal(size=p.row_ptr.size - 1)
    return vals, rhs


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_alone_equals_mix(hip_device, prec):
    pats = _patterns([1500, 2500, 4000])
    reqs = [(q % 3, 1000 + q) for q in range(12)]
    # alone: one slot, one request at a time
    alone = {}
    s1 = ca.Server(slots=1, prec=prec)
    ids = [s1.add_pattern(p.row_ptr, p.col, p.val, p.piv_starts, p.block_starts)
           for _, p in pats]
    for pi, seed in reqs:
        k, p = pats[pi]
        vals, rhs = _request(k, p, seed)
        r = s1.submit(ids[pi], vals, rhs)
        x, res, t, slot = s1.wait(r)
        alone[(pi, seed)] = x
        K = sp.csr_matrix((vals, p.col, p.row_ptr))
        assert np.linalg.norm(rhs - K @ x) / np.linalg.norm(rhs) <= 1e-8, res
    s1.close()
    # mix: four slots, everything submitted at once
    s4 = ca.Server(slots=4, prec=prec)
    ids = [s4.add_pattern(p.row_ptr, p.col, p.val, p.piv_starts, p.block_starts)
           for _, p in pats]
    handles = []
    for pi, seed in reqs:
        k, p = pats[pi]
        vals, rhs = _request(k, p, seed)
        handles.append(((pi, seed), s4.submit(ids[pi], vals, rhs)))
    slots = set()
    for key, r in handles:
        x, res, t, slot = s4.wait(r)
        slots.add(slot)
        assert np.array_equal(x, alone[key]), key
        assert res["true_rel_residual"] <= 1e-8
    assert len(slots) > 1, "the mix must have run on several slots"
    st = s4.stats()
    assert st["completed"] == len(reqs) and st["peak"] <= st["budget"]
    s4.close()


def test_server_equals_single_system_path(hip_device):
    import torch
    (k, p), = _patterns([3000])
    vals, rhs = _request(k, p, 7)
    s = ca.Server(slots=2, prec="f64")
    pid = s.add_pattern(p.row_ptr, p.col, p.val, p.piv_starts, p.block_starts)
    x_srv, res, _, _ = s.wait(s.submit(pid, vals, rhs))
    s.close()
    g = ca.BlockiLDLt(p.row_ptr, p.col, vals, p.piv_starts, p.block_starts, "f64")
    g.compute(0, 1.0, 1e-4)
    A = ca.CSRMatrixSPMV(p.row_ptr, p.col, vals)
    q = ca.SQMR(A, 1e-8, 1000, g, true_residual_period=20, max_restarts=10)
    x = torch.zeros(rhs.size, dtype=torch.float64, device=hip_device)
    q.solve(torch.tensor(rhs, device=hip_device), x)
    assert np.array_equal(x.cpu().numpy(), x_srv)
    assert q.result["iterations"] == res["iterations"]


def test_budget_admission(hip_device):
    pats = _patterns([2000, 3000])
    # learn the footprint of one instance of the larger pattern
    s = ca.Server(slots=1, prec="f64")
    ids = [s.add_pattern(p.row_ptr, p.col, p.val, p.piv_starts, p.block_starts)
           for _, p in pats]
    k, p = pats[1]
    x_ref, _, _, _ = s.wait(s.submit(ids[1], *_request(k, p, 3)))
    st = s.stats()
    s.close()
    # a budget with room for the patterns and about one instance: requests
    # on three slots are served one instance at a time, with the same results
    s = ca.Server(slots=3, prec="f64", hbm_budget=int(st["peak"] * 1.2))
    ids = [s.add_pattern(p.row_ptr, p.col, p.val, p.piv_starts, p.block_starts)
           for _, p in pats]
    hs = [s.submit(ids[q % 2], *_request(*pats[q % 2], 3)) for q in range(6)]
    outs = [s.wait(h) for h in hs]
    for q, (x, res, t, slot) in enumerate(outs):
        assert res["true_rel_residual"] <= 1e-8
        if q % 2 == 1:
            assert np.array_equal(x, x_ref)
    st2 = s.stats()
    assert st2["peak"] <= st2["budget"]
    s.close()
    # a budget below the static data of a pattern is refused at registration
    s = ca.Server(slots=1, prec="f64", hbm_budget=1 << 16)
    with pytest.raises(ca.CulipError) as e:
        s.add_pattern(p.row_ptr, p.col, p.val, p.piv_starts, p.block_starts)
    assert e.value.status == 6
    s.close()


def test_pattern_solve_schedule(hip_device):
    """A pattern served with block-Jacobi solves gives the single-system
    result under the same schedule, bit for bit; the schedule is fixed once a
    request for the pattern was submitted."""
    import torch
    (k, p), = _patterns([3000])
    vals, rhs = _request(k, p, 11)
    s = ca.Server(slots=2, prec="f32")
    pid = s.add_pattern(p.row_ptr, p.col, p.val, p.piv_starts, p.block_starts)
    s.set_pattern_solve(pid, "jacobi", 2)
    x_srv, res, _, _ = s.wait(s.submit(pid, vals, rhs))
    with pytest.raises(ca.CulipError) as e:
        s.set_pattern_solve(pid, "levels")
    assert e.value.status == 4
    s.close()
    assert res["true_rel_residual"] <= 1e-8
    g = ca.BlockiLDLt(p.row_ptr, p.col, vals, p.piv_starts, p.block_starts, "f32")
    g.compute(0, 1.0, 1e-4)
    g.set_solve_algorithm("jacobi", 2)
    A = ca.CSRMatrixSPMV(p.row_ptr, p.col, vals)
    q = ca.SQMR(A, 1e-8, 1000, g, true_residual_period=20, max_restarts=10)
    x = torch.zeros(rhs.size, dtype=torch.float64, device=hip_device)
    q.solve(torch.tensor(rhs, device=hip_device), x)
    assert np.array_equal(x.cpu().numpy(), x_srv)
    assert q.result["iterations"] == res["iterations"]
    # the server-wide default applies to every pattern
    s = ca.Server(slots=1, prec="f32", solve_algorithm="jacobi", jacobi_sweeps=2)
    pid = s.add_pattern(p.row_ptr, p.col, p.val, p.piv_starts, p.block_starts)
    x2, _, _, _ = s.wait(s.submit(pid, vals, rhs))
    s.close()
    assert np.array_equal(x2, x_srv)


def test_single_launch_solves_four_slots(hip_device):
    """Four requests at once on four slots: four single launches resident
    together, each equal to the one-system path bit for bit (more slots:
    test_gpu_concurrency)."""
    import torch
    (k, p), = _patterns([3000])
    reqs = [_request(k, p, 12 + r) for r in range(4)]
    s = ca.Server(slots=4, prec="f32", solve_algorithm="single")
    pid = s.add_pattern(p.row_ptr, p.col, p.val, p.piv_starts, p.block_starts)
    ids = [s.submit(pid, v, r) for v, r in reqs]
    outs = [s.wait(i) for i in ids]
    s.close()
    for (vals, rhs), (x_srv, res, _, _) in zip(reqs, outs):
        assert res["true_rel_residual"] <= 1e-8
        g = ca.BlockiLDLt(p.row_ptr, p.col, vals, p.piv_starts, p.block_starts, "f32")
        g.compute(0, 1.0, 1e-4)
        A = ca.CSRMatrixSPMV(p.row_ptr, p.col, vals)
        q = ca.SQMR(A, 1e-8, 1000, g, true_residual_period=20, max_restarts=10)
        x = torch.zeros(rhs.size, dtype=torch.float64, device=hip_device)
        q.solve(torch.tensor(rhs, device=hip_device), x)
        assert np.array_equal(x.cpu().numpy(), x_srv)
