"""Numeric pivoting on the device (BlockiLDLt<T, bk, rook>,
block_ildlt_instances.cu:54-60) against the oracle: Bunch-Kaufman and rook
pivot choices inside pivot blocks, symmetric interchanges, off-diagonal
rows and the columns above in pivoting order, P in the solves.

FP64: the factor (metadata, stores, D, row norms), the permutation and the
pivot structure equal the oracle's bit for bit; solves agree to 1e-11."""
import numpy as np
import pytest

import culip_amd as ca
import oracle_lib as ol
from conftest import EDGE_CASES, edge_problem, sym_problem

pytestmark = pytest.mark.gpu


def _pair(csr, piv, bs, mode, t="d", fill=0, ff=1.0, thr=1e-4, prec="f64"):
    rp, ci, cv = csr
    g = ca.BlockiLDLt(rp, ci, cv, piv, bs, "f32" if t == "f" else "f64")
    g.set_pivoting(mode)
    g.set_pivot_precision(prec)
    g.compute(fill, ff, thr)
    o = ol.OracleBildlt(rp, ci, cv, piv, bs, t)
    o.set_pivoting(mode)
    o.set_pivot_precision(prec)
    o.symbolic(fill, ff)
    o.factor(thr)
    return g, o


def _equal(g, o):
    for nm in ("fblocks", "fix_store", "perm", "fis_piv"):
        assert np.array_equal(g.get_int(nm), o.get(nm)), nm
    for nm in ("fval_store", "tri", "row_norms"):
        a, b = g.get_real(nm), o.get(nm)
        assert np.array_equal(a, b), (nm, np.abs(a - b).max())


@pytest.mark.parametrize("mode", ["bk", "rook"])
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_pivoting_factor_bitwise_fp64(hip_device, mode, seed):
    csr, piv, bs, _ = sym_problem(n_blocks=9, bmax=32, density=0.12, two_by_two=False,
                                  seed=seed, diag_shift=0.3)
    g, o = _pair(csr, piv, bs, mode, fill=1, ff=2.0, thr=1e-3)
    assert not np.array_equal(o.get("perm"), np.arange(csr[0].size - 1))
    _equal(g, o)


@pytest.mark.parametrize("mode", ["bk", "rook"])
def test_pivoting_complete_fill_and_solves(hip_device, mode):
    import torch
    csr, piv, bs, A = sym_problem(n_blocks=6, bmax=24, density=1.0, two_by_two=False,
                                  seed=8, diag_shift=0.2)
    g, o = _pair(csr, piv, bs, mode, thr=0.0)
    _equal(g, o)
    n = A.shape[0]
    b = np.random.default_rng(4).normal(size=n)
    db = torch.tensor(b, device=hip_device)
    for algo, sw in (("single", 0), ("levels", 0), ("jacobi", 64)):
        g.set_solve_algorithm(algo, sw)
        for side in ("left", "middle", "right"):
            x = torch.zeros_like(db)
            getattr(g, f"solve_{side}")(db, x)
            torch.cuda.synchronize()
            ref = o.solve(side, b)
            err = np.linalg.norm(x.cpu().numpy() - ref) / np.linalg.norm(ref)
            assert err < 1e-11, (algo, side, err)
    # with complete fill the preconditioner is A itself: SQMR converges at once
    g.set_solve_algorithm("single")
    Aop = ca.CSRMatrixSPMV(*csr)
    q = ca.SQMR(Aop, 1e-10, 200, g, true_residual_period=5, max_restarts=5)
    x = torch.zeros_like(db)
    q.solve(db, x)
    assert np.linalg.norm(b - A @ x.cpu().numpy()) / np.linalg.norm(b) < 1e-10
    assert q.result["iterations"] <= 10


def _pivoting_kkt():
    """n = 2,000 KKT system whose H diagonal spans 1e-4 .. 1e4 (log-uniform)
    with unit off-diagonals: some pivot blocks interchange rows and take 2x2
    pivots under both rules (46-51 rows move). H is indefinite and the
    incomplete factor does not make SQMR converge on it, with or without
    pivoting (the oracle's SQMR diverges too), so it serves the factor
    comparison only."""
    k = ca.KKT(2000, 800, 4, 3, h_diag=(1e-4, 1e4), h_diag_log=True, h_offdiag_std=1.0, seed=0)
    p = ca.Prep(k.row_ptr, k.col, k.val, 2000, 16, ca.ORDER_BLOCK_COLOR)
    return k, p


@pytest.mark.parametrize("mode", ["bk", "rook"])
def test_pivoting_kkt_factor_bitwise_fp64(hip_device, mode):
    k, p = _pivoting_kkt()
    g, o = _pair((p.row_ptr, p.col, p.val), p.piv_starts, p.block_starts, mode)
    assert (o.get("perm") != np.arange(p.row_ptr.size - 1)).sum() > 20
    _equal(g, o)


@pytest.mark.parametrize("mode", ["bk", "rook"])
def test_pivoting_kkt_sqmr(hip_device, mode):
    """The n=2,000 KKT system of the benchmark generator (H diagonal 1..10):
    its pivot blocks keep their order under both rules (the diagonal
    dominates), the FP64 factor equals the oracle's and SQMR (single-launch
    and Jacobi solves) meets 1e-8 through the pivoting path."""
    import torch
    import scipy.sparse as sp
    k = ca.KKT(2000, 800, 4, 3, seed=0)
    p = ca.Prep(k.row_ptr, k.col, k.val, 2000, 16, ca.ORDER_BLOCK_COLOR)
    g, o = _pair((p.row_ptr, p.col, p.val), p.piv_starts, p.block_starts, mode)
    _equal(g, o)
    A = ca.CSRMatrixSPMV(p.row_ptr, p.col, p.val)
    b = k.rhs[p.perm]
    db = torch.tensor(b, device=hip_device)
    K = sp.csr_matrix((p.val, p.col, p.row_ptr))
    for algo, sw in (("single", 0), ("jacobi", 2)):
        g.set_solve_algorithm(algo, sw)
        q = ca.SQMR(A, 1e-8, 2000, g, true_residual_period=20, max_restarts=10)
        x = torch.zeros_like(db)
        q.solve(db, x)
        rel = np.linalg.norm(b - K @ x.cpu().numpy()) / np.linalg.norm(b)
        assert rel <= 1e-8, (algo, rel, q.result)


@pytest.mark.parametrize("mode", ["bk", "rook"])
def test_pivoting_fp32_close_to_fp64(hip_device, mode):
    csr, piv, bs, A = sym_problem(n_blocks=6, bmax=24, density=1.0, two_by_two=False,
                                  seed=8, diag_shift=2.0)
    g, _ = _pair(csr, piv, bs, mode, t="f", thr=0.0)
    od = ol.OracleBildlt(*csr, piv, bs, "d")
    od.set_pivoting(mode)
    od.symbolic(0, 1.0)
    od.factor(0.0)
    if np.array_equal(g.get_int("perm"), od.get("perm")):
        a, b = g.get_real("fval_store"), od.get("fval_store")
        assert np.abs(a - b).max() <= 1e-3 * np.abs(b).max()


@pytest.mark.parametrize("prec", ["factor", "f64"])
@pytest.mark.parametrize("mode", ["bk", "rook"])
def test_pivoting_kkt_fp32_bitwise(hip_device, mode, prec):
    """FP32 factor with Bunch-Kaufman / rook pivoting on the pivoting KKT
    system equals the FP32 oracle bit for bit (pivot choices, interchanges
    and values), with the 2x2 arithmetic of BlockiLDLt<float, bk, rook>
    ('factor') and with the library's FP64 2x2 arithmetic ('f64')."""
    k, p = _pivoting_kkt()
    g, o = _pair((p.row_ptr, p.col, p.val), p.piv_starts, p.block_starts, mode, t="f",
                 prec=prec)
    assert (o.get("perm") != np.arange(p.row_ptr.size - 1)).sum() > 20
    _equal(g, o)


@pytest.mark.parametrize("mode", ["bk", "rook"])
@pytest.mark.parametrize("case", EDGE_CASES)
def test_pivoting_edge_cases(hip_device, case, mode):
    """The edge cases of the static path under numeric pivoting (a single
    row, a single 2x2 pivot, one-row blocks, a diagonal matrix, a full 32 x
    32 block where 14-21 rows move, zero pivots where 10-11 rows move): the
    factor, permutation and pivot structure equal the oracle's bit for bit
    and every solve schedule agrees with the oracle's solves, which apply
    P."""
    import torch
    csr, piv, bs, A = edge_problem(case)
    g, o = _pair(csr, piv, bs, mode, thr=1e-4)
    _equal(g, o)
    n = A.shape[0]
    b = np.random.default_rng(n).normal(size=n)
    db = torch.tensor(b, device=hip_device)
    for algo, sw in (("levels", 0), ("single", 0), ("jacobi", 64)):
        g.set_solve_algorithm(algo, sw)
        for side in ("left", "middle", "right"):
            x = torch.zeros_like(db)
            getattr(g, f"solve_{side}")(db, x)
            torch.cuda.synchronize()
            ref = o.solve(side, b)
            err = np.linalg.norm(x.cpu().numpy() - ref) / max(np.linalg.norm(ref), 1e-300)
            assert err < 1e-12, (algo, side, err)
