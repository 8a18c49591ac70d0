"""Single-launch triangular solves from many streams at once, and the
reporting of a timed-out dependency wait.

A single-launch solve needs all of its waves resident; the library admits
at most four at a time per device and runs the others as level launches
(the same bits). The error of a solve whose bounded wait gave up is forced
from the host here (the testing hook sets the flag the kernel would set),
so no kernel is made to wait in vain."""
import numpy as np
import pytest

import culip_amd as ca

pytestmark = pytest.mark.gpu


def _system(n, m, seed=0, B=16):
    k = ca.KKT(n, m, 4, 3, seed=seed)
    p = ca.Prep(k.row_ptr, k.col, k.val, n, B, ca.ORDER_BLOCK_COLOR)
    return k, p


def test_eight_streams_of_single_launch_solves(hip_device):
    """Eight objects solving on eight streams at the same time (twice the
    number of single launches that fit): every result equals the same
    object's level-launch solve, bit for bit, and no wait times out."""
    import torch
    k, p = _system(6000, 2400)
    gs = []
    for s in range(8):
        vals = p.val * (1.0 + 0.01 * s)
        g = ca.BlockiLDLt(p.row_ptr, p.col, vals, p.piv_starts, p.block_starts,
                          "f32" if s % 2 else "f64")
        g.compute(0, 1.0, 1e-4)
        gs.append(g)
    N = p.row_ptr.size - 1
    rng = np.random.default_rng(1)
    bs = [torch.tensor(rng.normal(size=N), device=hip_device) for _ in gs]
    # references: level launches, one object after the other
    ref = []
    for g, b in zip(gs, bs):
        g.set_solve_algorithm("levels")
        xs = []
        for side in ("right", "left"):
            x = b.clone()
            getattr(g, f"solve_{side}")(x, x)
            xs.append(x)
        torch.cuda.synchronize()
        ref.append([x.cpu().numpy() for x in xs])
    streams = [torch.cuda.Stream() for _ in gs]
    outs = [[b.clone() for _ in range(2)] for b in bs]
    torch.cuda.synchronize()
    for g in gs:
        g.set_solve_algorithm("single")
    for rep in range(3):
        for g, st, o, b in zip(gs, streams, outs, bs):
            with torch.cuda.stream(st):
                o[0].copy_(b)
                o[1].copy_(b)
                g.solve_right(o[0], o[0], stream=st.cuda_stream)
                g.solve_left(o[1], o[1], stream=st.cuda_stream)
        torch.cuda.synchronize()
        for i, (o, r) in enumerate(zip(outs, ref)):
            assert np.array_equal(o[0].cpu().numpy(), r[0]), (rep, i, "right")
            assert np.array_equal(o[1].cpu().numpy(), r[1]), (rep, i, "left")
    # nothing was recorded: a further solve on every object succeeds
    for g, b in zip(gs, bs):
        x = b.clone()
        g.solve_right(x, x)
    torch.cuda.synchronize()


def test_many_slots_single_launch_server(hip_device):
    """A server with 8 slots and single-launch solves serves 8 requests at
    once; each result equals the one-system path bit for bit."""
    import torch
    k, p = _system(3000, 1200)
    reqs = []
    for r in range(8):
        vals = p.permute_values(k.barrier_values(1e-2, 1e2, seed=40 + r))
        rhs = k.rhs[p.perm].copy()
        reqs.append((vals, rhs))
    s = ca.Server(slots=8, prec="f32", solve_algorithm="single")
    pid = s.add_pattern(p.row_ptr, p.col, p.val, p.piv_starts, p.block_starts)
    ids = [s.submit(pid, v, r) for v, r in reqs]
    outs = [s.wait(i) for i in ids]
    s.close()
    for (vals, rhs), (x_srv, res, _, _) in zip(reqs, outs):
        g = ca.BlockiLDLt(p.row_ptr, p.col, vals, p.piv_starts, p.block_starts, "f32")
        g.compute(0, 1.0, 1e-4)
        A = ca.CSRMatrixSPMV(p.row_ptr, p.col, vals)
        q = ca.SQMR(A, 1e-8, 1000, g, true_residual_period=20, max_restarts=10)
        x = torch.zeros(rhs.size, dtype=torch.float64, device=hip_device)
        q.solve(torch.tensor(rhs, device=hip_device), x)
        assert np.array_equal(x.cpu().numpy(), x_srv)
        assert q.result == res


def test_solve_error_reported_once(hip_device):
    """A recorded wait timeout is reported by exactly one later call on the
    object (a direct solve, a refactor or an SQMR solve) and cleared: the
    call after it succeeds."""
    import torch
    k, p = _system(2000, 800)
    g = ca.BlockiLDLt(p.row_ptr, p.col, p.val, p.piv_starts, p.block_starts, "f64")
    g.compute(0, 1.0, 1e-4)
    b = torch.tensor(k.rhs[p.perm], device=hip_device)
    x = torch.zeros_like(b)
    g._inject_solve_error()
    with pytest.raises(ca.CulipError) as e:
        g.solve_right(b, x)
    assert "timed out" in str(e.value)
    g.solve_right(b, x)
    torch.cuda.synchronize()
    g._inject_solve_error()
    with pytest.raises(ca.CulipError):
        g.refactor(torch.tensor(p.val, device=hip_device), 1e-4)
    g.refactor(torch.tensor(p.val, device=hip_device), 1e-4)
    A = ca.CSRMatrixSPMV(p.row_ptr, p.col, p.val)
    q = ca.SQMR(A, 1e-8, 1000, g, true_residual_period=20, max_restarts=10)
    g._inject_solve_error()
    with pytest.raises(ca.CulipError):
        q.solve(b, x)
    x.zero_()
    q.solve(b, x)
    assert q.result["true_rel_residual"] <= 1e-8
