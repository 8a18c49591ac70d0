"""Refactors and SQMR iterations replayed from captured HIP graphs give the
same bits as the launches themselves. Each graph is replayed several times
with new values, with eager work and synchronisations in between, and every
replay is checked against the eager result."""
import numpy as np
import pytest

import culip_amd as ca
import culip_amd.batch as cb

pytestmark = pytest.mark.gpu


def _factor_state(g):
    return [g.get_int("fblocks"), g.get_int("fix_store"), g.get_real("fval_store"),
            g.get_real("tri"), g.get_real("row_norms")]


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_refactor_graph_replays_equal_eager(hip_device, prec):
    import torch
    k, p = cb.kkt_system(6000, 21, 16)
    eager = ca.BlockiLDLt(p.row_ptr, p.col, p.val, p.piv_starts, p.block_starts, prec)
    graph = ca.BlockiLDLt(p.row_ptr, p.col, p.val, p.piv_starts, p.block_starts, prec)
    for g in (eager, graph):
        g.compute(0, 1.0, 1e-4)
    graph.set_graphs(True)
    st = torch.cuda.Stream()
    b = torch.tensor(np.random.default_rng(3).normal(size=p.row_ptr.size - 1),
                     device=hip_device)
    for rep in range(4):
        vals = p.permute_values(k.barrier_values(1e-3, 1e3, seed=50 + rep))
        dv = torch.tensor(vals, device=hip_device)
        torch.cuda.synchronize()
        eager.refactor(dv, 1e-4)
        # eager work between replays on the default stream
        junk = torch.randn(1 << 16, device=hip_device).sum()
        torch.cuda.synchronize()
        with torch.cuda.stream(st):
            graph.refactor(dv, 1e-4, stream=st.cuda_stream)
        torch.cuda.synchronize()
        for a, c in zip(_factor_state(eager), _factor_state(graph)):
            assert np.array_equal(a, c), rep
        for side in ("left", "right"):
            x1, x2 = torch.zeros_like(b), torch.zeros_like(b)
            getattr(eager, f"solve_{side}")(b, x1)
            getattr(graph, f"solve_{side}")(b, x2)
            torch.cuda.synchronize()
            assert np.array_equal(x1.cpu().numpy(), x2.cpu().numpy()), (rep, side)
        del junk


@pytest.mark.parametrize("algo,sw", [("jacobi", 2), ("single", 0)])
def test_server_graph_replays_equal_eager(hip_device, algo, sw):
    """Requests served with graph replays (refactor, and SQMR segments
    between true-residual checks) equal those served with launches, bit
    for bit, over several requests per pattern."""
    pats = [cb.kkt_system(n, 30 + j, 16) for j, n in enumerate((3000, 9000, 5000, 7000))]
    reqs = []
    for r in range(12):
        j = r % 4
        k, p = pats[j]
        vals = p.permute_values(k.barrier_values(1e-3, 1e3, seed=80 + r))
        reqs.append((j, vals, k.rhs[p.perm].copy()))
    outs = []
    for graphs, slots in ((False, 1), (True, 1), (True, 4)):
        # one slot: each request alone; four slots: captures of some slots
        # overlap instance creation and solves on the others
        s = ca.Server(slots=slots, prec="f32", solve_algorithm=algo, jacobi_sweeps=sw,
                      equilibrate_sweeps=8, max_it=3000, graphs=graphs)
        pids = [s.add_pattern(p.row_ptr, p.col, p.val, p.piv_starts, p.block_starts)
                for _, p in pats]
        ids = [s.submit(pids[j], v, b) for j, v, b in reqs]
        outs.append([s.wait(i) for i in ids])
        s.close()
    for q, ((xe, re, _, _), (xg, rg, _, _), (xc, rc, _, _)) in enumerate(zip(*outs)):
        assert np.array_equal(xe, xg) and np.array_equal(xe, xc), q
        assert re == rg == rc, q
        assert re["true_rel_residual"] <= 1e-8
