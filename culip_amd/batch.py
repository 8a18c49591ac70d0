"""Batches of independent KKT systems of uneven size: the block-size and
solve-schedule autotuner, and the batched solve over a request server.

The autotuner sorts systems into size buckets and, per bucket, times every
candidate (maximum block size, triangular-solve schedule) on the bucket's
first system: numeric refactor plus SQMR to the FP64 true-residual
tolerance, the median of a few repetitions after a warm-up. A candidate that
misses the tolerance is never picked.

The batch runs on culip_amd.Batch: every system registered under its
bucket's block size, all refactored concurrently and solved by one SQMR
recurrence whose steps are single launches over all systems still
iterating, with block-Jacobi triangular solves (one sweep count for the
batch). solve_batch runs the same systems on a Server instead (one request
per system, `slots` streams), the path with exact triangular solves.
"""
import time

import numpy as np

from . import (Batch, BlockiLDLt, CSRMatrixSPMV, KKT, ORDER_BLOCK_COLOR, Prep, SQMR, Server)

BUCKETS = ((1000, 10000), (10000, 20000), (20000, 30000), (30000, 40001))
BLOCK_SIZES = (8, 16, 32)


def bucket_of(n, buckets=BUCKETS):
    for i, (lo, hi) in enumerate(buckets):
        if lo <= n < hi:
            return i
    raise ValueError(f"system size {n} outside every bucket")


def schedules(slots):
    """Candidate triangular-solve schedules: level launches and block-Jacobi
    with 2 or 3 sweeps, and the single launch (which the library runs as
    level launches whenever four others are in flight)."""
    return [("single", 0), ("levels", 0), ("jacobi", 2), ("jacobi", 3)]


def kkt_system(n, seed, B, order=ORDER_BLOCK_COLOR, m=None):
    """The batch generator: n primal, m = 0.4 n dual, 4 H off-diagonals
    per row, 3 A entries per column; returns (KKT, Prep)."""
    k = KKT(n, int(0.4 * n) if m is None else m, 4, 3, seed=seed)
    p = Prep(k.row_ptr, k.col, k.val, n, B, order)
    return k, p


def _time_candidate(p, rhs, prec, B, algo, sweeps, reps, tol, dev):
    import torch
    g = BlockiLDLt(p.row_ptr, p.col, p.val, p.piv_starts, p.block_starts, prec)
    g.compute(0, 1.0, 1e-4)
    A = CSRMatrixSPMV(p.row_ptr, p.col, p.val)
    b = torch.tensor(rhs, device=dev)
    dv = torch.tensor(p.val, device=dev)
    g.set_solve_algorithm(algo, sweeps)
    q = SQMR(A, tol, 1000, g, true_residual_period=20, max_restarts=10)
    x = torch.zeros_like(b)
    g.refactor(dv, 1e-4)
    q.solve(b, x)
    torch.cuda.synchronize()
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        g.refactor(dv, 1e-4)
        q.solve(b, x)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    if q.result["true_rel_residual"] > tol:
        return float("inf")
    return float(np.median(times)) * 1e3


def autotune(sizes, make_system, prec="f32", slots=16, block_sizes=BLOCK_SIZES, reps=3,
             tol=1e-8, buckets=BUCKETS, log=None):
    """Per size bucket present in `sizes`: (block size, schedule, sweeps)
    with the shortest refactor + solve of the bucket's first system.
    make_system(i, B) -> (KKT, Prep) of system i with maximum block B.
    Returns (choice by bucket index, all timings in ms by key)."""
    import torch
    dev = torch.device("cuda:0")
    tuned, timings = {}, {}
    for bi, (lo, hi) in enumerate(buckets):
        idx = [i for i, n in enumerate(sizes) if lo <= n < hi]
        if not idx:
            continue
        i = idx[0]
        best = None
        for B in block_sizes:
            k, p = make_system(i, B)
            rhs = k.rhs[p.perm]
            for algo, sw in schedules(slots):
                ms = _time_candidate(p, rhs, prec, B, algo, sw, reps, tol, dev)
                name = algo if algo != "jacobi" else f"{algo}{sw}"
                timings[f"{lo}-{hi - 1}:B{B}:{name}"] = ms
                if best is None or ms < best[0]:
                    best = (ms, B, algo, sw)
        if best[0] == float("inf"):
            raise RuntimeError(f"autotune: no candidate meets the tolerance in bucket {lo}-{hi - 1}")
        tuned[bi] = best[1:]
        if log:
            log(f"autotune bucket {lo}-{hi - 1}: block size {best[1]}, {best[2]} {best[3]}")
    return tuned, timings


def autotune_batch(sizes, make_system, prec="f32", block_sizes=BLOCK_SIZES, sweep_options=(2, 3),
                   reps=5, tol=1e-8, buckets=BUCKETS, log=None):
    """For the batched engine: a block size per size bucket and one
    block-Jacobi sweep count for the whole batch. Every (block size, sweeps)
    pair is timed on each bucket's first system as in autotune; the sweep
    count is the one whose per-bucket best times sum lowest, and each
    bucket takes its fastest block size at that count.
    Returns (block size by bucket index, sweeps, all timings in ms)."""
    import torch
    dev = torch.device("cuda:0")
    per, timings, present = {}, {}, []
    for bi, (lo, hi) in enumerate(buckets):
        idx = [i for i, n in enumerate(sizes) if lo <= n < hi]
        if not idx:
            continue
        present.append(bi)
        for B in block_sizes:
            k, p = make_system(idx[0], B)
            rhs = k.rhs[p.perm]
            for sw in sweep_options:
                ms = _time_candidate(p, rhs, prec, B, "jacobi", sw, reps, tol, dev)
                per[(bi, B, sw)] = ms
                timings[f"{lo}-{hi - 1}:B{B}:jacobi{sw}"] = ms
    if not present:
        raise ValueError("autotune_batch: no system in any bucket")
    total = {sw: sum(min(per[(bi, B, sw)] for B in block_sizes) for bi in present)
             for sw in sweep_options}
    sweeps = min(sweep_options, key=lambda sw: total[sw])
    if total[sweeps] == float("inf"):
        raise RuntimeError("autotune_batch: no sweep count meets the tolerance in every bucket")
    tuned = {bi: min(block_sizes, key=lambda B: per[(bi, B, sweeps)]) for bi in present}
    if log:
        for bi in present:
            lo, hi = buckets[bi]
            log(f"autotune bucket {lo}-{hi - 1}: block size {tuned[bi]}, jacobi {sweeps}")
    return tuned, sweeps, timings


def solve_batch_engine(systems, prec, sweeps, tol=1e-8, max_it=1000, passes=2, log=None,
                       device_arrays=True):
    """Solves every system (prep, rhs) on one Batch. Registration (symbolic
    analysis and upload, 16 threads) is timed apart; each pass refactors
    and solves all systems; the first pass also allocates the batch's
    pooled work space and builds its launch lists. device_arrays: values,
    right-hand sides and solutions stay in device memory (copied there
    before the passes), else every pass takes them from host memory and
    returns x to it. Returns (results of the last pass as (x, result)
    per system, seconds per pass, registration seconds, Batch.info() after
    the last pass)."""
    import concurrent.futures as cf
    bt = Batch(prec=prec, tol=tol, max_it=max_it, jacobi_sweeps=sweeps)
    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(16) as ex:
        idx = list(ex.map(lambda s: bt.add(s[0].row_ptr, s[0].col, s[0].val, s[0].piv_starts,
                                           s[0].block_starts), systems))
    t_reg = time.perf_counter() - t0
    S = len(systems)
    vals, rhs = [None] * S, [None] * S
    for j, i in enumerate(idx):
        vals[i] = systems[j][0].val
        rhs[i] = systems[j][1]
    if device_arrays:
        import torch
        vals = [torch.tensor(np.ascontiguousarray(v), dtype=torch.float64, device="cuda:0")
                for v in vals]
        rhs = [torch.tensor(np.ascontiguousarray(r), dtype=torch.float64, device="cuda:0")
               for r in rhs]
        torch.cuda.synchronize()
    secs = []
    xs = res = None
    for rep in range(passes):
        t0 = time.perf_counter()
        xs, res = bt.solve(vals, rhs)
        secs.append(time.perf_counter() - t0)
        if log:
            inf = bt.info()
            log(f"batch pass {rep} (one batch, {prec}, jacobi {sweeps}): "
                f"{S / secs[-1]:.1f} solves/s (refactors {inf['factor_s'] * 1e3:.0f} ms, "
                f"SQMR {inf['sqmr_s'] * 1e3:.0f} ms)")
    phases = bt.info()
    bt.close()
    if device_arrays:
        xs = [x.cpu().numpy() for x in xs]
    return [(xs[idx[j]], res[idx[j]]) for j in range(S)], secs, t_reg, phases


def solve_batch(systems, schedule_of, slots, prec, tol=1e-8, max_it=1000, passes=2, log=None,
                graphs=False):
    """Solves every system (prep, rhs) on a Server with `slots` streams.
    schedule_of(i) -> (algorithm, sweeps) or None for the server default
    (single launch). Registration (symbolic analysis and upload) is timed
    apart; each pass submits every system at once and waits for all; the
    first pass also creates the slots' numeric instances. graphs: replay
    refactors and SQMR segments from captured HIP graphs.
    Returns (results of the last pass, seconds per pass, registration
    seconds)."""
    import concurrent.futures as cf
    srv = Server(slots=slots, prec=prec, tol=tol, max_it=max_it, solve_algorithm="single",
                 graphs=graphs)
    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(16) as ex:
        pids = list(ex.map(lambda s: srv.add_pattern(s[0].row_ptr, s[0].col, s[0].val,
                                                     s[0].piv_starts, s[0].block_starts),
                           systems))
    t_reg = time.perf_counter() - t0
    for i, pid in enumerate(pids):
        sch = schedule_of(i)
        if sch is not None:
            srv.set_pattern_solve(pid, sch[0], sch[1])
    secs = []
    out = None
    for rep in range(passes):
        t0 = time.perf_counter()
        ids = [srv.submit(pid, s[0].val, s[1]) for pid, s in zip(pids, systems)]
        out = [srv.wait(r) for r in ids]
        secs.append(time.perf_counter() - t0)
        if log:
            log(f"batch pass {rep} ({slots} slots, {prec}): {len(out) / secs[-1]:.1f} solves/s")
    srv.close()
    return out, secs, t_reg


def single_system_solve(p, rhs, prec, algo, sweeps, tol=1e-8, max_it=1000):
    """The one-system path with the server's SQMR options: x (host) and the
    result dict."""
    import torch
    g = BlockiLDLt(p.row_ptr, p.col, p.val, p.piv_starts, p.block_starts, prec)
    g.compute(0, 1.0, 1e-4)
    g.set_solve_algorithm(algo, sweeps)
    A = CSRMatrixSPMV(p.row_ptr, p.col, p.val)
    q = SQMR(A, tol, max_it, g, true_residual_period=20, max_restarts=10)
    b = torch.tensor(np.ascontiguousarray(rhs), device="cuda:0")
    x = torch.zeros_like(b)
    q.solve(b, x)
    return x.cpu().numpy(), dict(q.result)
