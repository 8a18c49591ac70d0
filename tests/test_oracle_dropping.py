"""Pins of the oracle's dual dropping on over-capacity sparse blocks
(k_apply_block_dual_dropping, block_ildlt_kernels.impl.cuh:1866-2074):
threshold drop, borrowing from the leftover pool (:1948-1978) and the
secondary drop to the largest entries (CUB radix sort, :2012-2061).

The matrices are built so that the off-diagonal block (1, 2) before dropping
is known in closed form: block 0 is the identity, so U_01 = A_01 and
U_02 = A_02; A_11 = I + A_01' A_01, so block 1's pivot factor is the identity
again, and the scaled block is T = A_12 - A_01' A_02, computed here in exact
integer (or dyadic) arithmetic. The kept entries are then checked against a
numpy brute force: the `keep` largest |T| with ties to the lower tile index
(row-major within a 32 x 32 tile).
"""
import math

import numpy as np
import pytest

import oracle_lib as ol

B = 32
NZ_EPS = 1e-10


def _csr(A):
    n = A.shape[0]
    rp, ci, cv = [0], [], []
    for i in range(n):
        nz = np.nonzero(A[i])[0]
        ci += list(nz)
        cv += list(A[i, nz])
        rp.append(len(ci))
    return np.array(rp, np.int32), np.array(ci, np.int32), np.array(cv)


def _sparse_int(rng, nnz, vals=(-2, -1, 1, 2)):
    M = np.zeros((B, B))
    pos = rng.choice(B * B, size=nnz, replace=False)
    M.flat[pos] = rng.choice(vals, size=nnz)
    return M


def _overflow_problem(seed, nnz01, nnz02, nnz12, small12=0):
    """Three 32-row blocks; returns (A, T) with T the scaled block (1, 2)
    before dropping and A's entry count in block (1, 2)."""
    rng = np.random.default_rng(seed)
    A01 = _sparse_int(rng, nnz01)
    A02 = _sparse_int(rng, nnz02)
    A12 = _sparse_int(rng, nnz12)
    if small12:
        # dyadic small entries at positions A_01' A_02 leaves zero, so that a
        # threshold can drop them while every sum stays exact
        P = A01.T @ A02
        free = np.flatnonzero((P == 0) & (A12 == 0))
        pos = rng.choice(free, size=small12, replace=False)
        A12.flat[pos] = rng.choice([0.125, -0.125, 0.0625], size=small12)
    n = 3 * B
    A = np.zeros((n, n))
    A[0:B, 0:B] = np.eye(B)
    A[0:B, B:2 * B] = A01
    A[0:B, 2 * B:] = A02
    A[B:2 * B, B:2 * B] = np.eye(B) + A01.T @ A01
    A[B:2 * B, 2 * B:] = A12
    A[2 * B:, 2 * B:] = 1000.0 * np.eye(B)
    A = np.triu(A) + np.triu(A, 1).T
    T = A12 - A01.T @ A02
    return A, T, int(np.count_nonzero(A12)), int(np.count_nonzero(A01)), \
        int(np.count_nonzero(A02))


def _run(A, ff, thr):
    rp, ci, cv = _csr(A)
    o = ol.OracleBildlt(rp, ci, cv, np.arange(A.shape[0]), [0, B, 2 * B], "d")
    o.symbolic(0, ff)
    o.factor(thr)
    return o


def _block(o, br, bc):
    fb = o.get("fblocks").reshape(-1, 9)
    sb = o.get("blocks").reshape(-1, 9)
    i = np.flatnonzero((fb[:, 5] == br) & (fb[:, 6] == bc))[0]
    return fb[i], sb[i]


def _brute_force_keep(T, keep, thr_per_row=None):
    """Entries of T that survive: |v| > NZ_EPS and not below the row's
    threshold, then the `keep` largest |v|, ties to the lower tile index."""
    cand = []
    for r in range(B):
        for c in range(B):
            v = T[r, c]
            if abs(v) <= NZ_EPS:
                continue
            if thr_per_row is not None and abs(v) < thr_per_row[r]:
                continue
            cand.append((-abs(v), 32 * r + c, r * B + c, v))
    cand.sort()
    return cand[:keep], len(cand)


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_over_capacity_keeps_largest_with_index_ties(seed):
    """Fill factor 1: no leftover exists, the block keeps exactly its
    capacity (A's entry count) of its largest entries."""
    A, T, nnz12, _, _ = _overflow_problem(seed, 40, 40, 6)
    o = _run(A, 1.0, 0.0)
    f, s = _block(o, 1, 2)
    assert f[1] == 0, "block (1, 2) must be sparse"
    assert s[3] == nnz12 and f[4] == 0
    want, survivors = _brute_force_keep(T, nnz12)
    assert survivors > nnz12, "case must overflow"
    # equal magnitudes inside the kept range or across its cut, or the tie
    # rule would not be exercised
    allc, _ = _brute_force_keep(T, B * B)
    mags = [w[0] for w in want]
    assert len(set(mags)) < len(mags) or allc[nnz12 - 1][0] == allc[nnz12][0]
    assert f[2] == nnz12
    fix, fv = o.get("fix_store"), o.get("fval_store")
    np.testing.assert_array_equal(fix[f[7]:f[7] + nnz12], [w[2] for w in want])
    np.testing.assert_array_equal(fv[f[8]:f[8] + nnz12], [w[3] for w in want])


def _check_borrowing(seed):
    """Fill factor 1.5: the level-0 sparse blocks (0, 1) and (0, 2) release
    ceil(1.5 nnz) - nnz each into the pool; the overflowing block (1, 2)
    receives min(request, pool) extra entries. A partial grant keeps that
    many largest entries (largest first); a full grant leaves the block
    within its raised capacity, stored row-major like any other."""
    ff = 1.5
    A, T, nnz12, nnz01, nnz02 = _overflow_problem(seed, 30, 30, 4)
    o = _run(A, ff, 0.0)
    f, s = _block(o, 1, 2)
    cap = math.ceil(ff * nnz12)
    assert s[3] == cap
    pool = (math.ceil(ff * nnz01) - nnz01) + (math.ceil(ff * nnz02) - nnz02)
    _, survivors = _brute_force_keep(T, B * B)
    assert survivors > cap + 1
    granted = min(survivors - cap, pool)
    assert granted > 0
    assert f[4] == granted
    keep = cap + granted
    assert f[2] == keep and f[3] == keep
    want, _ = _brute_force_keep(T, keep)
    if keep == survivors:
        want.sort(key=lambda w: w[1])     # row-major
    fix, fv = o.get("fix_store"), o.get("fval_store")
    np.testing.assert_array_equal(fix[f[7]:f[7] + keep], [w[2] for w in want])
    np.testing.assert_array_equal(fv[f[8]:f[8] + keep], [w[3] for w in want])
    return keep < survivors


@pytest.mark.parametrize("seed", [4, 5, 6])
def test_borrowing_from_the_leftover_pool(seed):
    _check_borrowing(seed)


def test_borrowing_cases_cover_partial_and_full_grants():
    partial = [_check_borrowing(s) for s in (4, 5, 6)]
    assert any(partial) and not all(partial)


def test_threshold_drop_before_capacity():
    """Small entries fall to the threshold first (|v| < tau sqrtf(row norm),
    the norm summing the squares of the row's U entries, unit diagonal
    included); the capacity then applies to what is left."""
    A, T, nnz12, _, _ = _overflow_problem(9, 40, 40, 6, small12=10)
    tau = 0.03
    rn = 1.0 + (T * T).sum(axis=1)   # integer and dyadic: exact in any order
    thr = tau * np.sqrt(rn.astype(np.float32)).astype(np.float64)
    small = (np.abs(T) > NZ_EPS) & (np.abs(T) < thr[:, None])
    assert small.any(), "case must exercise the threshold drop"
    o = _run(A, 1.0, tau)
    np.testing.assert_array_equal(o.get("row_norms")[B:2 * B], rn)
    f, _ = _block(o, 1, 2)
    want, survivors = _brute_force_keep(T, nnz12, thr)
    assert survivors > nnz12
    fix, fv = o.get("fix_store"), o.get("fval_store")
    np.testing.assert_array_equal(fix[f[7]:f[7] + nnz12], [w[2] for w in want])
    np.testing.assert_array_equal(fv[f[8]:f[8] + nnz12], [w[3] for w in want])


def test_under_capacity_stores_row_major():
    """A block within capacity stores every surviving entry in row-major
    order (d_save_block, :85-173)."""
    A, T, nnz12, _, _ = _overflow_problem(10, 8, 8, 30)
    o = _run(A, 4.0, 0.0)
    f, s = _block(o, 1, 2)
    nz = [(r * B + c, T[r, c]) for r in range(B) for c in range(B) if abs(T[r, c]) > NZ_EPS]
    assert len(nz) <= s[3]
    assert f[2] == len(nz) and f[4] == 0
    fix, fv = o.get("fix_store"), o.get("fval_store")
    np.testing.assert_array_equal(fix[f[7]:f[7] + f[2]], [q for q, _ in nz])
    np.testing.assert_array_equal(fv[f[8]:f[8] + f[2]], [v for _, v in nz])


@pytest.mark.parametrize("seed,ff,thr", [(7, 1.0, 0.05), (8, 1.0, 0.2), (8, 1.3, 0.01),
                                         (11, 1.0, 0.0), (12, 2.0, 0.1)])
def test_pool_invariants(seed, ff, thr):
    """Over every sparse block in level-list order (the order the pool is
    drawn in): a grant never exceeds what earlier blocks released, so the
    stored entries of every prefix of blocks fit in that prefix's capacity
    (what makes the reference's in-place compaction safe); the pool before
    a block is the capacity its predecessors left unused; a granted block is
    full; the final pool is the unused sparse capacity."""
    from conftest import sym_problem
    csr, piv, bs, _ = sym_problem(n_blocks=12, bmax=24, density=0.1, two_by_two=False,
                                  seed=seed)
    o = ol.OracleBildlt(*csr, piv, bs, "d")
    o.symbolic(1, ff)
    o.factor(thr)
    fb = o.get("fblocks").reshape(-1, 9)
    sb = o.get("blocks").reshape(-1, 9)
    order = o.get("lvlb_idx")
    sparse = [b for b in order if sb[b, 1] == 0]
    assert sparse
    cap_prefix = stored_prefix = 0
    any_grant = False
    for b in sparse:
        cap0, granted, stored = sb[b, 3], fb[b, 4], fb[b, 2]
        assert granted >= 0
        assert stored <= cap0 + granted
        assert fb[b, 3] == cap0 + granted
        # the grant comes out of space earlier blocks left unused
        assert granted <= cap_prefix - stored_prefix
        if granted > 0:
            any_grant = True
            assert stored == cap0 + granted
        cap_prefix += cap0
        stored_prefix += stored
        assert stored_prefix <= cap_prefix
    sc = o.get("scalars")
    assert sc[13] == sum(int(fb[b, 2]) for b in sparse)     # out_ix
    assert sc[13] <= sc[11]                                  # within ix_space
    assert sc[15] == sc[11] - sc[13]                         # final pool
    if seed in (7, 8) and ff == 1.0:
        assert any_grant, "case must exercise borrowing"
