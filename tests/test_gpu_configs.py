"""GPU tests at the shapes of the benchmark workloads.

* kkt-40k with the kkt-400k generator parameters (6 H off-diagonals per
  row, 4 A entries per column, maximum block 32, colour ordering): FP64
  factor bit for bit against the oracle, solves against the oracle's.
* kkt-400k itself (n = 400,000, m = 150,000): SQMR to a true relative
  residual of 1e-8, single-launch and level-launch runs bit-identical.
* A mini batch spanning the four size buckets with the autotuner: every
  system equals the one-system path bit for bit, at 1e-8.
* Serving over two patterns of 50,000 and 60,000 unknowns: every request's
  result is the same alone and in the mix.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import culip_amd as ca
import culip_amd.batch as cb
import oracle_lib as ol

pytestmark = pytest.mark.gpu


def _kkt(n, m, o, c, B, seed=0, **kw):
    k = ca.KKT(n, m, o, c, seed=seed, **kw)
    p = ca.Prep(k.row_ptr, k.col, k.val, n, B, ca.ORDER_BLOCK_COLOR)
    return k, p


def test_kkt40k_generator_of_kkt400k_fp64_parity(hip_device):
    import torch
    k, p = _kkt(40000, 15000, 6, 4, 32)
    g = ca.BlockiLDLt(p.row_ptr, p.col, p.val, p.piv_starts, p.block_starts, "f64")
    g.compute(0, 1.0, 1e-4)
    o = ol.OracleBildlt(p.row_ptr, p.col, p.val, p.piv_starts, p.block_starts, "d")
    o.symbolic(0, 1.0)
    o.factor(1e-4)
    for nm in ("blocks", "lvlb_idx", "coarse_csr_off", "coarse_csr_col"):
        assert np.array_equal(g.get_int(nm), o.get(nm)), nm
    assert np.array_equal(g.get_int("fblocks"), o.get("fblocks"))
    assert np.array_equal(g.get_int("fix_store"), o.get("fix_store"))
    for nm in ("fval_store", "tri", "row_norms"):
        assert np.array_equal(g.get_real(nm), o.get(nm)), nm
    rng = np.random.default_rng(40)
    b = rng.normal(size=p.row_ptr.size - 1)
    db = torch.tensor(b, device=hip_device)
    for side in ("left", "middle", "right"):
        x = torch.zeros_like(db)
        getattr(g, f"solve_{side}")(db, x)
        torch.cuda.synchronize()
        ref = o.solve(side, b)
        err = np.linalg.norm(x.cpu().numpy() - ref) / np.linalg.norm(ref)
        assert err < 1e-11, (side, err)


def test_kkt400k_solves_to_tolerance(hip_device):
    import torch
    k, p = _kkt(400000, 150000, 6, 4, 32)
    g = ca.BlockiLDLt(p.row_ptr, p.col, p.val, p.piv_starts, p.block_starts, "f64")
    g.compute(0, 1.0, 1e-4)
    A = ca.CSRMatrixSPMV(p.row_ptr, p.col, p.val)
    b_h = k.rhs[p.perm]
    b = torch.tensor(b_h, device=hip_device)
    out = {}
    for algo in ("single", "levels"):
        g.set_solve_algorithm(algo)
        q = ca.SQMR(A, 1e-8, 1000, g, true_residual_period=20, max_restarts=10)
        x = torch.zeros_like(b)
        q.solve(b, x)
        torch.cuda.synchronize()
        out[algo] = (x.cpu().numpy(), dict(q.result))
    xs, rs = out["single"]
    xl, rl = out["levels"]
    assert np.array_equal(xs, xl) and rs == rl
    K = sp.csr_matrix((p.val, p.col, p.row_ptr))
    rel = np.linalg.norm(b_h - K @ xs) / np.linalg.norm(b_h)
    assert rel <= 1e-8 and rs["true_rel_residual"] <= 1e-8, (rel, rs)


def test_mini_batch_autotuned(hip_device):
    sizes = [1500, 6000, 12000, 17000, 22000, 27000, 31000, 38000,
             2500, 9000, 14000, 19500, 21000, 29000, 33000, 36000]

    def make(i, B):
        return cb.kkt_system(sizes[i], 100 + i, B)

    tuned, timings = cb.autotune(sizes, make, prec="f32", slots=8, reps=1)
    assert sorted(tuned) == [0, 1, 2, 3]
    assert any(np.isfinite(v) for v in timings.values())
    systems = []
    for i, n in enumerate(sizes):
        B, algo, sw = tuned[cb.bucket_of(n)]
        k, p = make(i, B)
        systems.append((p, k.rhs[p.perm].copy()))
    res, secs, _ = cb.solve_batch(systems, lambda i: tuned[cb.bucket_of(sizes[i])][1:], slots=8,
                                  prec="f32", passes=1)
    for i, ((p, rhs), (x_b, r_b, _, _)) in enumerate(zip(systems, res)):
        _, algo, sw = tuned[cb.bucket_of(sizes[i])]
        x_1, r_1 = cb.single_system_solve(p, rhs, "f32", algo, sw)
        assert np.array_equal(x_b, x_1), i
        assert r_b == r_1, i
        assert r_b["true_rel_residual"] <= 1e-8, (i, r_b)


def test_serve_two_large_patterns_alone_equals_mix(hip_device):
    pats = [cb.kkt_system(n, 500 + j, 32) for j, n in enumerate((50000, 60000))]
    reqs = []
    for j, (k, p) in enumerate(pats):
        for r in range(3):
            vals = p.permute_values(k.barrier_values(1e-4, 1e4, seed=70 + 10 * j + r))
            reqs.append((j, vals, k.rhs[p.perm].copy()))
    opts = dict(prec="f32", max_it=3000, solve_algorithm="jacobi", jacobi_sweeps=2)
    # in the mix: every request submitted at once on 4 slots
    s = ca.Server(slots=4, **opts)
    pids = [s.add_pattern(p.row_ptr, p.col, p.val, p.piv_starts, p.block_starts)
            for _, p in pats]
    ids = [s.submit(pids[j], v, b) for j, v, b in reqs]
    mix = [s.wait(i) for i in ids]
    s.close()
    # alone: one request at a time on one slot
    s = ca.Server(slots=1, **opts)
    pids = [s.add_pattern(p.row_ptr, p.col, p.val, p.piv_starts, p.block_starts)
            for _, p in pats]
    alone = [s.wait(s.submit(pids[j], v, b)) for j, v, b in reqs]
    s.close()
    for q, ((x_m, r_m, _, _), (x_a, r_a, _, _)) in enumerate(zip(mix, alone)):
        assert np.array_equal(x_m, x_a), q
        assert r_m == r_a, q
    for (j, vals, rhs), (x, r, _, _) in zip(reqs, mix):
        p = pats[j][1]
        K = sp.csr_matrix((vals, p.col, p.row_ptr))
        rel = np.linalg.norm(rhs - K @ x) / np.linalg.norm(rhs)
        assert abs(rel - r["true_rel_residual"]) <= 1e-9 + 1e-6 * rel
