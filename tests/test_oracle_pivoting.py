"""Pins of the oracle's numeric pivoting (BlockiLDLt<T, bk, rook>,
block_ildlt_instances.cu:54-60): Bunch-Kaufman and rook pivot selection
inside pivot blocks (block_ildlt_kernels.impl.cuh:743-948), symmetric
interchanges (:679-739), the row order of off-diagonal blocks (:1733-1751)
and the column order of the blocks above (:1290-1418), and P in the solves.

* On a single block the Bunch-Kaufman factor is LAPACK's (scipy.linalg.ldl,
  ?SYTRF): the same permutation, the same 1x1 / 2x2 pivot structure and
  the same D.
* With complete block fill (every block dense, nothing dropped) the block
  factor is a complete LDL' of the permuted matrix: A(p, p) = U' D U, which
  pins the interchanges of off-diagonal rows and of the columns above.
* Rook pivoting keeps every multiplier below 1 / (1 - alpha).
"""
import numpy as np
import pytest
import scipy.linalg as sl

import oracle_lib as ol
from conftest import sym_problem
from test_oracle_math import _factor_dense, _to_csr

ALPHA = np.float32((1.0 + np.sqrt(np.float32(17.0))) / 8.0)


def _oracle(csr, piv, bs, t, mode, fill=0, ff=1.0, thr=0.0):
    o = ol.OracleBildlt(*csr, piv, bs, t)
    o.set_pivoting(mode)
    o.symbolic(fill, ff)
    o.factor(thr)
    return o


def _dense_D(o, n):
    tri = o.get("tri")
    D = np.zeros((n, n))
    for i in range(n):
        D[i, i] = tri[3 * i + 1]
        if i + 1 < n:
            D[i + 1, i] = tri[3 * i + 2]
            D[i, i + 1] = tri[3 * (i + 1)]
    return D


@pytest.mark.parametrize("seed", range(12))
def test_bunch_kaufman_single_block_equals_lapack(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(5, 33))
    A = rng.normal(size=(n, n))
    A = A + A.T
    o = _oracle(_to_csr(A), np.arange(n), [0], "d", "bk")
    perm, fpiv = o.get("perm"), o.get("fis_piv")
    lu, d, sperm = sl.ldl(A, lower=True)
    assert np.array_equal(perm, sperm)
    two = [i for i in range(n - 1) if d[i + 1, i] != 0]
    assert [i for i in range(n - 1) if fpiv[i + 1] == 0] == two
    np.testing.assert_allclose(o.get("tri")[1::3], np.diag(d), rtol=1e-10)
    U, _ = _factor_dense(o, n)
    D = _dense_D(o, n)
    np.testing.assert_allclose(U.T @ D @ U, A[np.ix_(perm, perm)], rtol=0,
                               atol=1e-12 * np.abs(A).max())


@pytest.mark.parametrize("seed", range(8))
def test_rook_single_block_bounded_multipliers(seed):
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(8, 33))
    A = rng.normal(size=(n, n)) * 10.0 ** rng.uniform(-3, 3, size=(n, n))
    A = A + A.T
    o = _oracle(_to_csr(A), np.arange(n), [0], "d", "rook")
    perm = o.get("perm")
    assert np.array_equal(np.sort(perm), np.arange(n))
    U, _ = _factor_dense(o, n)
    D = _dense_D(o, n)
    np.testing.assert_allclose(U.T @ D @ U, A[np.ix_(perm, perm)], rtol=0,
                               atol=1e-11 * np.abs(A).max())
    off = U - np.diag(np.diag(U))
    # 2x2 pivot blocks of U are the identity; every multiplier is bounded
    assert np.abs(off).max() <= 1.0 / (1.0 - ALPHA) + 1e-9


@pytest.mark.parametrize("mode", ["bk", "rook"])
@pytest.mark.parametrize("t", ["d", "f"])
def test_complete_block_fill_with_pivoting_is_exact(mode, t):
    csr, piv, bs, A = sym_problem(n_blocks=5, bmax=12, density=1.0, two_by_two=False,
                                  seed=21, diag_shift=0.2)
    n = A.shape[0]
    o = _oracle(csr, piv, bs, t, mode, thr=1e-4)
    fb = o.get("fblocks").reshape(-1, 9)
    assert (fb[:, 1] == 1).all()
    perm = o.get("perm")
    assert not np.array_equal(perm, np.arange(n)), "case must interchange rows"
    # interchanges stay inside their pivot block
    for a, b in zip(list(bs), list(bs[1:]) + [n]):
        assert sorted(perm[a:b]) == list(range(a, b))
    U, _ = _factor_dense(o, n)
    D = _dense_D(o, n)
    tol = 1e-9 if t == "d" else 2e-3
    np.testing.assert_allclose(U.T @ D @ U, A[np.ix_(perm, perm)], rtol=tol,
                               atol=tol * np.abs(A).max())


@pytest.mark.parametrize("mode", ["bk", "rook"])
def test_solves_apply_the_permutation(mode):
    csr, piv, bs, A = sym_problem(n_blocks=8, bmax=20, density=0.15, seed=5, diag_shift=0.3)
    n = A.shape[0]
    o = _oracle(csr, piv, bs, "d", mode, fill=1, ff=3.0, thr=1e-4)
    perm = o.get("perm")
    U, _ = _factor_dense(o, n)
    D = _dense_D(o, n)
    b = np.random.default_rng(2).normal(size=n)
    xr = o.solve("right", b)              # x[perm] = U \\ b
    assert np.linalg.norm(U @ xr[perm] - b) / np.linalg.norm(b) < 1e-10 * np.linalg.cond(U)
    xl = o.solve("left", b)               # x = U' \\ b[perm]
    assert np.linalg.norm(U.T @ xl - b[perm]) / np.linalg.norm(b) < 1e-10 * np.linalg.cond(U)
    xm = o.solve("middle", b)
    assert np.linalg.norm(D @ xm - b) / np.linalg.norm(b) < 1e-12 * np.linalg.cond(D)


def test_static_mode_is_unchanged():
    csr, piv, bs, A = sym_problem(n_blocks=6, bmax=16, density=0.2, seed=3)
    o1 = _oracle(csr, piv, bs, "d", "static", fill=1, ff=2.0, thr=1e-3)
    o2 = ol.OracleBildlt(*csr, piv, bs, "d")
    o2.symbolic(1, 2.0)
    o2.factor(1e-3)
    for nm in ("fblocks", "fix_store", "fval_store", "tri", "row_norms"):
        assert np.array_equal(o1.get(nm), o2.get(nm)), nm
    assert np.array_equal(o1.get("perm"), np.arange(A.shape[0]))
    assert np.array_equal(o1.get("fis_piv"), o1.get("is_piv"))


def _block_values(o):
    """(metadata rows, {block: stored values}) of a factored oracle."""
    fb = o.get("fblocks").reshape(-1, 9)
    bs = np.asarray(o.get("block_starts"), np.int64)
    fv = o.get("fval_store")
    out = {}
    for b in range(fb.shape[0]):
        h = bs[fb[b, 5] + 1] - bs[fb[b, 5]]
        w = bs[fb[b, 6] + 1] - bs[fb[b, 6]]
        cnt = int(h * w) if fb[b, 1] == 1 else int(fb[b, 2])
        out[b] = fv[fb[b, 8]:fb[b, 8] + cnt]
    return fb, out


@pytest.mark.parametrize("seed", [2, 3, 6, 11])
def test_pivot_precision_f64_tracks_the_fp64_factor(seed):
    """The oracle's FP64 arithmetic for the 2x2 steps and D solves of an
    FP32 factor (the library's FP32 default; the reference computes them in
    FP32). On random indefinite matrices with 32-45 static 2x2 pivots and
    no dropping, every block of that FP32 factor is within 1e-5 of the FP64
    factor (relative to the block's largest value) and at most twice as far
    from it as the FP32-arithmetic factor, plus 1e-6 (both are ~1e-6: FP32
    rounding elsewhere dominates). A dropped term or a wrong sign in the
    FP64 form leaves errors of order one."""
    csr, piv, bs, _ = sym_problem(n_blocks=9, bmax=32, density=0.12, seed=seed, diag_shift=20.0)
    assert np.diff(np.append(piv, csr[0].size - 1)).max() == 2
    ref = _oracle(csr, piv, bs, "d", "static")
    errs = {}
    for prec in ("factor", "f64"):
        o = ol.OracleBildlt(*csr, piv, bs, "f")
        o.set_pivot_precision(prec)
        o.symbolic(0, 1.0)
        o.factor(0.0)
        fb, vals = _block_values(o)
        fb_r, vals_r = _block_values(ref)
        assert np.array_equal(fb[:, :3], fb_r[:, :3])
        worst = 0.0
        for b, v in vals.items():
            scale = np.abs(vals_r[b]).max() if vals_r[b].size else 0.0
            if scale > 0:
                worst = max(worst, float(np.abs(v - vals_r[b]).max() / scale))
        errs[prec] = worst
    assert errs["f64"] < 1e-5, errs
    assert errs["f64"] <= 2.0 * errs["factor"] + 1e-6, errs


def _first_pivot_rows_exact(seed):
    """{precision: whether the first 2x2 pivot's rows of block row 0 equal
    the FP64 factor's rounded to FP32}; asserts it for 'f64'."""
    rng = np.random.default_rng(seed)
    n0, n1 = 6, 5
    n = n0 + n1
    A = rng.normal(size=(n, n)).astype(np.float32).astype(np.float64)
    A = np.triu(A, 1)
    A = A + A.T
    A[np.arange(n), np.arange(n)] = rng.normal(size=n).astype(np.float32) * 0.5
    piv = np.array([0, 2, 3, 5, 6, 7, 8, 9, 10], np.int32)   # 2x2 at rows 0-1 and 3-4
    bs = np.array([0, n0], np.int32)
    csr = _to_csr(A)
    od = _oracle(csr, piv, bs, "d", "static")
    fb, vd = _block_values(od)
    exact = {}
    for prec in ("f64", "factor"):
        o = ol.OracleBildlt(*csr, piv, bs, "f")
        o.set_pivot_precision(prec)
        o.symbolic(0, 1.0)
        o.factor(0.0)
        fbf, vf = _block_values(o)
        assert np.array_equal(fbf, fb)
        same = True
        for b in range(fb.shape[0]):
            if fb[b, 5] != 0:
                continue
            w = n0 if fb[b, 6] == 0 else n1
            assert fb[b, 1] == 1          # dense: row-major h x w
            rows_f = vf[b][:2 * w]
            rows_d = vd[b][:2 * w].astype(np.float32)
            if prec == "f64":
                assert np.array_equal(rows_f, rows_d.astype(np.float64)), (b, rows_f - rows_d)
            same = same and np.array_equal(rows_f, rows_d.astype(np.float64))
        exact[prec] = same
    return exact


@pytest.mark.parametrize("seed", range(6))
def test_pivot_precision_f64_first_pivot_is_the_rounded_fp64_step(seed):
    """Exact pin of the FP64 2x2 arithmetic. Matrix entries are FP32 values;
    block row 0 has no Schur updates and starts with a 2x2 pivot, so the FP32
    factor with FP64 2x2 arithmetic holds, in that pivot's two rows of U_00
    and of the off-diagonal block U_01, the FP64 factor's values rounded to
    FP32: the same double operations on the same inputs, rounded once. The
    reference's FP32 arithmetic rounds every step and differs from them on
    most draws (checked over the seeds)."""
    assert _first_pivot_rows_exact(seed)["f64"]


def test_pivot_precision_pin_discriminates():
    """The exact pin above fails for the FP32-arithmetic form on most draws:
    it is a property of the FP64 arithmetic, not of any FP32 factor."""
    differs = sum(not _first_pivot_rows_exact(s)["factor"] for s in range(6))
    assert differs >= 3, differs


def test_pivot_precision_changes_nothing_in_fp64():
    csr, piv, bs, _ = sym_problem(n_blocks=9, bmax=32, density=0.12, seed=3)
    outs = []
    for prec in ("factor", "f64"):
        o = ol.OracleBildlt(*csr, piv, bs, "d")
        o.set_pivot_precision(prec)
        o.symbolic(1, 2.0)
        o.factor(1e-3)
        outs.append([o.get(nm) for nm in ("fblocks", "fix_store", "fval_store", "tri", "row_norms")])
    for a, b in zip(*outs):
        assert np.array_equal(a, b)
