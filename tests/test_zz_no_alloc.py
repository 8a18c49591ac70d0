"""culip_bildlt_solve_noalloc (the allocation-free solve) against
culip_bildlt_solve, both through the C-ABI: the same bits for every side,
transpose flag and solve schedule, on two systems in a row (the second
smaller, so the shared static scratch holds the first call's contents), in
place and out of place; sizes above the scratch array are refused."""
import numpy as np
import pytest

import culip_amd as ca

pytestmark = pytest.mark.gpu

MAX_ROWS = 1 << 20


def _system(n, prec, algo, sweeps, seed):
    k = ca.KKT(n, int(0.4 * n), 4, 3, seed=seed)
    p = ca.Prep(k.row_ptr, k.col, k.val, n, 16, ca.ORDER_BLOCK_COLOR)
    g = ca.BlockiLDLt(p.row_ptr, p.col, p.val, p.piv_starts, p.block_starts, prec)
    g.compute(0, 1.0, 1e-4)
    g.set_solve_algorithm(algo, sweeps)
    return g, p.row_ptr.size - 1


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("algo,sweeps", [("jacobi", 2), ("jacobi", 3), ("single", 0)])
def test_noalloc_solve_equals_solve(hip_device, prec, algo, sweeps):
    import torch
    for n, seed in ((20000, 1), (3000, 2)):
        g, N = _system(n, prec, algo, sweeps, seed)
        rng = np.random.default_rng(seed)
        b = torch.tensor(rng.normal(size=N), device=hip_device)
        for side in ("left", "middle", "right"):
            for transpose in (False, True):
                ref = torch.zeros_like(b)
                getattr(g, f"solve_{side}")(b, ref, transpose=transpose)
                out = torch.full_like(b, float("nan"))
                g.solve_noalloc(side, b, out, transpose=transpose)
                inplace = b.clone()
                g.solve_noalloc(side, inplace, inplace, transpose=transpose)
                torch.cuda.synchronize()
                r = ref.cpu().numpy()
                assert np.isfinite(r).all()
                assert np.array_equal(out.cpu().numpy(), r), (n, side, transpose)
                assert np.array_equal(inplace.cpu().numpy(), r), (n, side, transpose, "in place")


@pytest.mark.parametrize("mode", ["bk", "rook"])
@pytest.mark.parametrize("algo,sweeps", [("jacobi", 2), ("single", 0)])
def test_noalloc_pivoted_equals_solve(hip_device, mode, algo, sweeps):
    """With numeric pivoting the solves pull P before the left solve and
    push it after the right one; the allocation-free form does this in the
    static scratch. Two pivoted systems in a row, the second smaller: the
    n = 2,000 KKT system with H diagonal 1e-4 .. 1e4 (rows move) and a
    complete-fill random indefinite system."""
    import torch
    from conftest import sym_problem
    k = ca.KKT(2000, 800, 4, 3, h_diag=(1e-4, 1e4), h_diag_log=True, h_offdiag_std=1.0, seed=0)
    p = ca.Prep(k.row_ptr, k.col, k.val, 2000, 16, ca.ORDER_BLOCK_COLOR)
    csr2, piv2, bs2, _ = sym_problem(n_blocks=6, bmax=24, density=1.0, two_by_two=False, seed=8,
                                     diag_shift=0.2)
    for (rp, ci, cv), piv, bs, seed in (((p.row_ptr, p.col, p.val), p.piv_starts,
                                         p.block_starts, 1), (csr2, piv2, bs2, 2)):
        g = ca.BlockiLDLt(rp, ci, cv, piv, bs, "f64")
        g.set_pivoting(mode)
        g.compute(0, 1.0, 1e-4)
        g.set_solve_algorithm(algo, sweeps)
        N = rp.size - 1
        assert (g.get_int("perm") != np.arange(N)).any()
        b = torch.tensor(np.random.default_rng(seed).normal(size=N), device=hip_device)
        for side in ("left", "middle", "right"):
            for transpose in (False, True):
                ref = torch.zeros_like(b)
                getattr(g, f"solve_{side}")(b, ref, transpose=transpose)
                out = torch.full_like(b, float("nan"))
                g.solve_noalloc(side, b, out, transpose=transpose)
                inplace = b.clone()
                g.solve_noalloc(side, inplace, inplace, transpose=transpose)
                torch.cuda.synchronize()
                r = ref.cpu().numpy()
                assert np.isfinite(r).all()
                assert np.array_equal(out.cpu().numpy(), r), (N, side, transpose)
                assert np.array_equal(inplace.cpu().numpy(), r), (N, side, transpose, "in place")


def test_noalloc_refuses_oversize(hip_device):
    import torch
    N = MAX_ROWS + 1
    rp = np.arange(N + 1, dtype=np.int32)
    ci = np.arange(N, dtype=np.int32)
    g = ca.BlockiLDLt(rp, ci, np.ones(N), np.arange(N), np.arange(0, N, 32), "f64")
    b = torch.zeros(1, dtype=torch.float64, device=hip_device)
    with pytest.raises(ca.CulipError) as e:
        g.solve_noalloc("right", b, b)
    assert e.value.status == 1 and "exceed" in str(e.value)
