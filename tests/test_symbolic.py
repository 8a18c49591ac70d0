"""Symbolic phase: the product's host implementation against the oracle
(bit for bit, every structure) and the oracle against brute-force
definitions of the coarse pattern, level-of-fill, level sets, level lists,
storage layout and format rule on small inputs."""
import math

import numpy as np
import pytest

import culip_amd as ca
import oracle_lib as ol
from conftest import EDGE_CASES, edge_problem, sym_problem

NAMES = ["block_starts", "coarse_csr_off", "coarse_csr_col", "coarse_csc_off", "coarse_csc_idx",
         "a_block_nnz", "lvl_off", "lvl_rows", "row_level", "lvlb_off", "lvlb_idx",
         "block_in_lvl", "row_first_block_offset", "blocks", "ix_store", "solve_off",
         "solve_idx", "is_piv"]


def _both(csr, piv, bs, t, fl, ff):
    g = ca.BlockiLDLt(*csr, piv, bs, "f32" if t == "f" else "f64")
    g.symbolic(fl, ff)
    o = ol.OracleBildlt(*csr, piv, bs, t)
    o.symbolic(fl, ff)
    return g, o


def _kkt_case(n, m, B, order, seed=0, **kw):
    k = ca.KKT(n, m, 4, 3, seed=seed, **kw)
    p = ca.Prep(k.row_ptr, k.col, k.val, n, B, order)
    return (p.row_ptr, p.col, p.val), p.piv_starts, p.block_starts


CASES = [
    ("kkt2k-natural", lambda: _kkt_case(2000, 800, 16, ca.ORDER_NATURAL), 0, 1.0),
    ("kkt2k-colour", lambda: _kkt_case(2000, 800, 16, ca.ORDER_BLOCK_COLOR), 0, 1.0),
    ("kkt2k-fill1", lambda: _kkt_case(2000, 800, 16, ca.ORDER_NATURAL), 1, 2.0),
    ("kkt-uneven-blocks", lambda: _kkt_case(1003, 397, 13, ca.ORDER_BLOCK_COLOR, seed=3), 0, 1.5),
    ("sym-2x2-fill2", lambda: sym_problem(n_blocks=14, bmax=32, density=0.08, seed=9)[:3], 2, 1.3),
    ("sym-dense", lambda: sym_problem(n_blocks=6, bmax=16, density=0.9, seed=2)[:3], 0, 1.0),
] + [(e, (lambda e=e: edge_problem(e)[:3]), 0, 1.0) for e in EDGE_CASES]


# thousands of block rows: the host threads split every per-row pass (the
# oracle's own brute-force checks stay on the small cases)
LARGE_CASES = [
    ("kkt40k-colour", lambda: _kkt_case(40000, 15000, 16, ca.ORDER_BLOCK_COLOR), 0, 1.0),
    ("kkt12k-fill1", lambda: _kkt_case(12000, 4800, 16, ca.ORDER_NATURAL), 1, 1.5),
]


@pytest.mark.parametrize("name,make,fl,ff", CASES + LARGE_CASES,
                         ids=[c[0] for c in CASES + LARGE_CASES])
@pytest.mark.parametrize("t", ["d", "f"])
def test_product_symbolic_bit_exact(name, make, fl, ff, t):
    csr, piv, bs = make()
    g, o = _both(csr, piv, bs, t, fl, ff)
    for nm in NAMES:
        a, e = g.get_int(nm), o.get(nm)
        assert a.shape == e.shape and np.array_equal(a, e), nm
    assert np.array_equal(g.get_real("val_store"), o.get("val_store"))
    sg, so = g.get_int("scalars"), o.get("scalars")
    assert np.array_equal(sg[:13], so[:13])


def _brute(csr, piv, bs, fill_level, t):
    rp, ci, cv = csr
    n = rp.size - 1
    bsx = list(bs) + [n]
    nb = len(bs)
    blk = np.zeros(n, int)
    for b in range(nb):
        blk[bsx[b]:bsx[b + 1]] = b
    adj = [set() for _ in range(nb)]
    upper = set()
    cnt = {}
    for r in range(n):
        for q in range(rp[r], rp[r + 1]):
            c = ci[q]
            adj[blk[r]].add(blk[c])
            if r <= c:
                upper.add((blk[r], blk[c]))
                cnt[(blk[r], blk[c])] = cnt.get((blk[r], blk[c]), 0) + 1
    pattern = set(upper)
    # level-l fill: k > i at graph distance 2..l+1 through nodes < i
    for i in range(nb):
        dist = {i: 0}
        frontier = [i]
        for step in range(1, fill_level + 2):
            nxt = []
            for u in frontier:
                if u != i and u > i:
                    continue
                for v in adj[u]:
                    if v not in dist:
                        dist[v] = step
                        if v < i and step <= fill_level:
                            nxt.append(v)
                        elif v > i and step >= 2:
                            pattern.add((i, v))
            frontier = nxt
        pattern.add((i, i))
    return pattern, cnt, bsx, nb


@pytest.mark.parametrize("name,make,fl,ff", CASES[2:], ids=[c[0] for c in CASES[2:]])
def test_oracle_symbolic_brute_force(name, make, fl, ff):
    csr, piv, bs = make()
    o = ol.OracleBildlt(*csr, piv, bs, "d")
    o.symbolic(fl, ff)
    pattern, cnt, bsx, nb = _brute(csr, piv, bs, fl, "d")
    off, col = o.get("coarse_csr_off"), o.get("coarse_csr_col")
    got = {(i, int(col[q])) for i in range(nb) for q in range(off[i], off[i + 1])}
    assert got == pattern
    ids = {(i, int(col[q])): q for i in range(nb) for q in range(off[i], off[i + 1])}
    # levels: longest path in the block DAG
    lev = np.zeros(nb, int)
    for (i, j) in sorted(pattern, key=lambda p: p[1]):
        pass
    for j in range(nb):
        preds = [i for (i, jj) in pattern if jj == j and i < j]
        lev[j] = 1 + max(lev[i] for i in preds) if preds else 0
    assert np.array_equal(o.get("row_level"), lev)
    lo = o.get("lvl_off")
    rows = o.get("lvl_rows")
    assert sorted(rows) == list(range(nb))
    for l in range(len(lo) - 1):
        assert set(rows[lo[l]:lo[l + 1]]) == set(np.nonzero(lev == l)[0])
    # level list: by (row level, pivot first, block id); offsets by counts
    meta = o.get("blocks").reshape(-1, 9)
    key = [(lev[m[5]], 0 if m[5] == m[6] else 1, b) for b, m in enumerate(meta)]
    assert list(o.get("lvlb_idx")) == sorted(range(len(meta)), key=lambda b: key[b])
    # storage: prefix sums of per-block space in level-list order
    sv = 8
    pos_ix, pos_val = 0, 0
    for b in o.get("lvlb_idx"):
        m = meta[b]
        r = bsx[m[5] + 1] - bsx[m[5]]
        c = bsx[m[6] + 1] - bsx[m[6]]
        nnz = cnt.get((m[5], m[6]), 0)
        assert m[2] == nnz
        if nnz == 0:
            row = [cnt.get((m[5], j), 0) for j in range(nb) if (m[5], j) in pattern]
            src = math.ceil(sum(row) / sum(1 for v in row if v))
        else:
            src = nnz
        cap = min(r * c, math.ceil(ff * src))
        assert m[3] == cap
        sparse = m[5] != m[6] and m[5] < 0.9 * nb and (4 + sv) * cap < 0.5 * sv * r * c
        assert m[1] == (0 if sparse else 1)
        assert m[7] == pos_ix and m[8] == pos_val
        pos_ix += cap if sparse else 0
        pos_val += cap if sparse else r * c
    # the Schur pairs: common k < i of columns i and j, ascending k
    g = ca.BlockiLDLt(*csr, piv, bs)
    g.symbolic(fl, ff)
    uo, uki, ukj = g.get_int("upd_off"), g.get_int("upd_ki"), g.get_int("upd_kj")
    for b, m in enumerate(meta):
        i, j = m[5], m[6]
        ks = [k for k in range(i) if (k, i) in pattern and (k, j) in pattern] if i != j else []
        assert [meta[x][5] for x in uki[uo[b]:uo[b + 1]]] == ks
        assert [ids[(k, i)] for k in ks] == list(uki[uo[b]:uo[b + 1]])
        assert [ids[(k, j)] for k in ks] == list(ukj[uo[b]:uo[b + 1]])


def test_format_threshold_depends_on_T():
    """nnz < r c / 3 sparse for double, nnz < r c / 4 for float
    (block_ildlt.impl.cuh:1027-1036)."""
    expect = {("d", 255): 0, ("d", 256): 0, ("d", 341): 0, ("d", 342): 1,
              ("f", 255): 0, ("f", 256): 1, ("f", 341): 1, ("f", 342): 1}
    for (t, k), fmt in expect.items():
        # block (0, 1) of four 32-row blocks holds k entries
        n = 128
        A = np.eye(n) * 4.0
        idx = np.random.default_rng(k).choice(32 * 32, size=k, replace=False)
        for q in idx:
            r, c = divmod(int(q), 32)
            A[r, 32 + c] = A[32 + c, r] = 0.01
        rp, ci, cv = [0], [], []
        for i in range(n):
            nz = np.nonzero(A[i])[0]
            ci += list(nz)
            cv += list(A[i, nz])
            rp.append(len(ci))
        args = (np.array(rp), np.array(ci), np.array(cv), np.arange(n), [0, 32, 64, 96])
        o = ol.OracleBildlt(*args, t)
        o.symbolic(0, 1.0)
        b01 = [m for m in o.get("blocks").reshape(-1, 9) if m[5] == 0 and m[6] == 1][0]
        assert b01[3] == k and b01[1] == fmt, (t, k, b01)
        g = ca.BlockiLDLt(*args, "f32" if t == "f" else "f64")
        g.symbolic(0, 1.0)
        assert np.array_equal(g.get_int("blocks"), o.get("blocks"))


def test_bildlt_rejects_empty_and_straddling_inputs():
    """No pivot starts (an empty matrix) and a 2x2 pivot straddling a block
    boundary (block starts must be pivot starts, quirk Q7 of the reference)
    are refused with INVALID."""
    csr, piv, bs, _ = edge_problem("one-2x2")
    with pytest.raises(ca.CulipError) as e:
        ca.BlockiLDLt(*csr, np.zeros(0, np.int32), bs)
    assert e.value.status == 1
    csr, piv, bs, _ = edge_problem("dense-32")
    g = ca.BlockiLDLt(*csr, piv, np.array([0, 1], np.int32))   # row 1 is inside a 2x2 pivot
    with pytest.raises(ca.CulipError) as e:
        g.symbolic(0, 1.0)
    assert e.value.status == 1
