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
