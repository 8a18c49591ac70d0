"""KKT solves/sec on one MI355X: BILDLt refactor + SQMR to true relative
residual <= 1e-8, with p50/p99 solve latency.

A step is one interior-point KKT solve: numeric refactor of the block
incomplete LDL' from the device-resident values of K (the pattern's symbolic
analysis is cached), then preconditioned SQMR from x = 0 until the FP64 true
relative residual ||b - K x|| / ||b|| is <= 1e-8 (refinement restarts when
the recurrence drifts). Inputs are resident in HBM when the timed region
starts. Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

WORKLOADS = {
    # Plain method at scale: n=400k, m=150k, 6 H off-diagonals/row, 4 A nnz/col,
    # max block 32, level-0 fill, drop tol 1e-4, SQMR to 1e-8, all FP64.
    "kkt-400k": dict(n=400000, m=150000, o=6, c=4, B=32, h_diag=(1.0, 10.0), log=False,
                     delta=1e-8),
    # CPU-runnable case: n=2,000, m=800, 4 off-diagonals/row, 3 nnz/col, block 16.
    "kkt-2k": dict(n=2000, m=800, o=4, c=3, B=16, h_diag=(1.0, 10.0), log=False, delta=1e-8),
    # Barrier-conditioned case: H = H0 + diag(w), w = 10^U(-6,6), delta 1e-10.
    "kkt-2k-barrier": dict(n=2000, m=800, o=4, c=3, B=16, h_diag=(1.0, 10.0), log=False,
                           delta=1e-10, barrier=(1e-6, 1e6)),
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="kkt-400k",
                    choices=sorted(WORKLOADS) + ["serve", "batch"])
    ap.add_argument("--prec", default=None, choices=["f64", "f32"],
                    help="factor precision (default f64; f32 for serve and batch)")
    # serve / batch
    ap.add_argument("--slots", type=int, default=0,
                    help="concurrent streams (default: 8 for serve, 16 for batch, the best "
                         "measured of 8/12/16)")
    ap.add_argument("--patterns", type=int, default=64)
    ap.add_argument("--n-lo", type=int, default=5000)
    ap.add_argument("--n-hi", type=int, default=200000)
    ap.add_argument("--requests", type=int, default=1000)
    ap.add_argument("--rate", type=float, default=0.0,
                    help="offered requests/s (0: 85%% of the warm-up throughput)")
    ap.add_argument("--max-block", type=int, default=32)
    ap.add_argument("--systems", type=int, default=512)
    ap.add_argument("--serve-jacobi-sweeps", type=int, default=2,
                    help="serve: block-Jacobi solves with this many sweeps (0: level solves)")
    ap.add_argument("--graphs", type=int, default=-1,
                    help="replay refactors (and, when serving, SQMR segments) from HIP graphs "
                         "(default: on for serve and the single-system workloads, off for "
                         "batch, where the small systems measured slower)")
    ap.add_argument("--equilibrate-sweeps", type=int, default=8,
                    help="serve: device equilibration sweeps per request (0: off)")
    ap.add_argument("--max-it", type=int, default=0,
                    help="SQMR iteration budget per request (0: 3000 for serve)")
    ap.add_argument("--order", default="color", choices=["color", "natural"])
    ap.add_argument("--fill-factor", type=float, default=1.0)
    ap.add_argument("--threshold", type=float, default=1e-4)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--plain-requests", type=int, default=64,
                    help="serve workload: requests of the stream served by the plain FP64 "
                         "per-system path for the speedup (0: skip)")
    ap.add_argument("--no-plain", action="store_true",
                    help="batch workload: skip the plain FP64 per-system comparison "
                         "(profiling runs)")
    ap.add_argument("--no-variants", action="store_true",
                    help="skip the FP32-factor and block-Jacobi sub-runs")
    ap.add_argument("--solve-algorithm", default="single",
                    choices=["single", "levels", "jacobi"],
                    help="triangular-solve schedule (set_solve_algorithm)")
    ap.add_argument("--jacobi-sweeps", type=int, default=0)
    ap.add_argument("--pmc-probe", action="store_true",
                    help="one step and one calibration copy, for rocprofv3 --pmc runs")
    return ap.parse_args()


PROBE_COPY_DOUBLES = 1 << 25   # 256 MiB read + 256 MiB written

# HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE (one run each, `--pmc-probe`), scaled by what the same counters
# read for the probe's device copy of known size (reads x2.000, writes
# x1.000); kept in profiles/kkt400k_hbm_traffic.md. Configurations without
# such a run report null.
MEASURED_TRAFFIC = {
    ("kkt-400k", "f64", "color", "single"): {"k_trsv_chain": 140.7e6, "k_csr_spmv": 233.1e6,
                                              "k_trsv_chain_product": 634.7e6},
    ("kkt-400k", "f64", "color", "levels"): {"k_trsv_level": 322.1e6, "k_csr_spmv": 269.8e6},
}


def event_time(fn, reps, stream=None):
    """Average device duration of fn() in ms, HIP events on torch's stream."""
    import torch
    fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def _cpu_info():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(), "threads": 1}


def _oracle_solve_times(wl, n, m, order, seed, fill_factor, threshold, repeats):
    """The oracle (the C restatement of the reference algorithm, one core)
    on the workload's generator at n, m: symbolic once (cached per pattern,
    as on the device path), then `repeats` timed solves of refactor + SQMR
    with the reference's stop and refinement rounds until the FP64 true
    relative residual is <= 1e-8 (the device path's stopping rule)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import culip_amd as ca
    import oracle_lib as ol
    kkt = ca.KKT(n, m, wl["o"], wl["c"], h_diag=wl["h_diag"], h_diag_log=wl["log"],
                 delta=wl["delta"], seed=seed, barrier=wl.get("barrier"))
    prep = ca.Prep(kkt.row_ptr, kkt.col, kkt.val, n, wl["B"], order)
    b = kkt.rhs[prep.perm]
    o = ol.OracleBildlt(prep.row_ptr, prep.col, prep.val, prep.piv_starts, prep.block_starts,
                        "d")
    t0 = time.perf_counter()
    o.symbolic(0, fill_factor)
    t_sym = time.perf_counter() - t0
    times, fac = [], []
    its = rounds = 0
    rel = None
    for _ in range(repeats):
        t0 = time.perf_counter()
        o.factor(threshold)
        t1 = time.perf_counter()
        x, its, rounds, rel = ol.sqmr_bildlt_refined(prep.row_ptr, prep.col, prep.val, o, b,
                                                     1e-8, 1000)
        t2 = time.perf_counter()
        times.append(t2 - t0)
        fac.append(t1 - t0)
    return {"times": times, "factor": fac, "symbolic": t_sym, "iterations": its,
            "rounds": rounds, "true_rel": rel, "nnz": int(prep.col.size), "x": x, "prep": prep,
            "oracle": o, "rhs": b}


def cpu_baseline(wl, order, seed, fill_factor, threshold):
    """The reference CPU path (the oracle, one core) on the same generator.
    For kkt-400k a bounded sample: the n=40,000 system of the same
    generator (1/10 of the nnz), 3 timed solves, scaled linearly by nnz to
    the workload (about 30 s of CPU work); smaller workloads are timed
    directly, 7 solves. Also the n=2,000 system (the reference's
    CPU-runnable case) timed directly, 7 solves, median and p99."""
    scale = 10 if wl["n"] >= 100000 else 1
    reps = 3 if scale > 1 else 7
    r = _oracle_solve_times(wl, wl["n"] // scale, wl["m"] // scale, order, seed, fill_factor,
                            threshold, reps)
    nnz_full = wl["n"] * (1 + wl["o"]) + 2 * wl["c"] * wl["n"] + wl["m"]
    ratio = nnz_full / r["nnz"]
    med = float(np.median(r["times"])) * ratio
    out = {"value": 1.0 / med, "unit": "solves/s", "cores": 1, "kind": "port",
           "ms_per_solve_median": med * 1e3,
           "ms_per_solve_p99": float(np.percentile(r["times"], 99)) * ratio * 1e3,
           "repeats": reps, "extrapolated": scale > 1,
           "sample": (f"oracle refactor + SQMR to FP64 true rel. residual <= 1e-8 (reference "
                      f"stop + {r['rounds']} refinement rounds, {r['iterations']} iterations, "
                      f"true rel {r['true_rel']:.1e}) on the same generator at "
                      f"n={wl['n'] // scale}, m={wl['m'] // scale} (seed {seed}), {reps} "
                      f"solves, median {np.median(r['times']):.2f}s (factor "
                      f"{np.median(r['factor']):.2f}s)"
                      + (f", scaled by the nnz ratio {ratio:.2f} to the workload" if scale > 1
                         else "")),
           "symbolic_s_once_per_pattern": r["symbolic"] * ratio}
    out.update(_cpu_info())
    if wl["n"] != 2000:
        small = WORKLOADS["kkt-2k"]
        q = _oracle_solve_times(small, small["n"], small["m"], order, seed, fill_factor,
                                threshold, 7)
        out["kkt_2k_direct"] = {
            "solves_per_s": 1.0 / float(np.median(q["times"])),
            "ms_per_solve_median": float(np.median(q["times"])) * 1e3,
            "ms_per_solve_p99": float(np.percentile(q["times"], 99)) * 1e3,
            "iterations": q["iterations"], "refinement_rounds": q["rounds"],
            "true_rel_residual": q["true_rel"], "repeats": 7,
            "symbolic_ms": q["symbolic"] * 1e3}
    return out


def precision_variant(order, seed, fill_factor, threshold):
    """The barrier-conditioned n=2,000 system with an FP32 factor (FP64
    SQMR, true residual every 20 iterations): error of its solution against
    the oracle's FP64 solve (reference BILDLt<double> + SQMR refined to
    1e-8) and against a sparse direct solve of K (scipy SuperLU)."""
    import scipy.sparse as sps
    import scipy.sparse.linalg as spl
    import torch
    import culip_amd as ca
    wl = WORKLOADS["kkt-2k-barrier"]
    r = _oracle_solve_times(wl, wl["n"], wl["m"], order, seed, fill_factor, threshold, 1)
    prep, b = r["prep"], r["rhs"]
    dev = torch.device("cuda:0")
    g = ca.BlockiLDLt(prep.row_ptr, prep.col, prep.val, prep.piv_starts, prep.block_starts,
                      "f32")
    g.compute(0, fill_factor, threshold)
    A = ca.CSRMatrixSPMV(prep.row_ptr, prep.col, prep.val)
    q = ca.SQMR(A, 1e-8, 3000, g, true_residual_period=20, max_restarts=20)
    db = torch.tensor(b, device=dev)
    x = torch.zeros_like(db)
    dv = torch.tensor(prep.val, device=dev)
    steps = 10
    q.solve(db, x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        g.refactor(dv, threshold)
        q.solve(db, x)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    xg = x.cpu().numpy()
    K = sps.csr_matrix((prep.val, prep.col, prep.row_ptr))
    xd = spl.spsolve(K.tocsc(), b)
    xo = r["x"]
    return {"workload": "kkt-2k-barrier (H = H0 + diag(10^U(-6,6)), delta 1e-10)",
            "prec": "f32 factor / f64 sqmr", "solves_per_s": 1.0 / dt,
            "sqmr_iterations": q.result["iterations"],
            "refinement_restarts": q.result["restarts"],
            "true_rel_residual": q.result["true_rel_residual"],
            "rel_error_vs_oracle_fp64_solve": float(np.linalg.norm(xg - xo) / np.linalg.norm(xo)),
            "oracle_fp64_true_rel_residual": r["true_rel"],
            "rel_error_vs_direct_solve": float(np.linalg.norm(xg - xd) / np.linalg.norm(xd)),
            "oracle_rel_error_vs_direct_solve":
                float(np.linalg.norm(xo - xd) / np.linalg.norm(xd))}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _pattern(ca, n, seed, B, order):
    k = ca.KKT(n, int(0.4 * n), 4, 3, seed=seed)
    p = ca.Prep(k.row_ptr, k.col, k.val, n, B, order)
    return k, p


def bench_serve(a):
    """Poisson stream of refactor + solve requests over resident patterns
    (n ~ U[5k, 200k], m = 0.4 n); each request rescales the barrier part of
    the H diagonal (H_ii = H0_ii + w_i, w_i log-uniform in [1e-4, 1e4])."""
    import concurrent.futures as cf
    import threading
    import culip_amd as ca
    rng = np.random.default_rng(a.seed)
    sizes = rng.integers(a.n_lo, a.n_hi + 1, size=a.patterns)
    order = ca.ORDER_BLOCK_COLOR
    sweeps = a.serve_jacobi_sweeps
    exact = "single"   # beyond four at a time the library runs level launches
    srv = ca.Server(slots=a.slots, prec=a.prec, max_it=a.max_it or 3000,
                    solve_algorithm="jacobi" if sweeps > 0 else exact, jacobi_sweeps=sweeps,
                    equilibrate_sweeps=a.equilibrate_sweeps, graphs=bool(a.graphs))
    nsets = 2

    def make(i):
        k, p = _pattern(ca, int(sizes[i]), 1000 + i, a.max_block, order)
        pid = srv.add_pattern(p.row_ptr, p.col, p.val, p.piv_starts, p.block_starts)
        vals = [p.permute_values(k.barrier_values(1e-4, 1e4, seed=10 * i + s))
                for s in range(nsets)]
        return pid, vals, k.rhs[p.perm].copy(), p.col.size

    t_setup = time.perf_counter()
    with cf.ThreadPoolExecutor(16) as ex:
        pats = list(ex.map(make, range(a.patterns)))
    t_setup = time.perf_counter() - t_setup
    log(f"{len(pats)} patterns registered in {t_setup:.1f}s")
    picks = rng.integers(0, a.patterns, size=a.requests + 2 * a.patterns)
    sets = rng.integers(0, nsets, size=picks.size)

    def run(order_ix, gaps):
        ids = [None] * len(order_ix)

        def feeder():
            t = time.perf_counter()
            for q, j in enumerate(order_ix):
                if gaps is not None:
                    t += gaps[q]
                    d = t - time.perf_counter()
                    if d > 0:
                        time.sleep(d)
                pid, vals, rhs, _ = pats[picks[j]]
                ids[q] = srv.submit(pid, vals[sets[j]], rhs)

        th = threading.Thread(target=feeder)
        t0 = time.perf_counter()
        th.start()
        out = []
        for q in range(len(order_ix)):
            while ids[q] is None:
                time.sleep(0.0005)
            out.append(srv.wait(ids[q]))
            if (q + 1) % 100 == 0:
                log(f"{q + 1} requests done")
        th.join()
        return out, time.perf_counter() - t0

    # warm-up: every pattern once per slot set-up, untimed; its throughput
    # calibrates the offered load of the timed stream
    warm, t_warm = run(list(range(a.requests, a.requests + 2 * a.patterns)), None)
    thr = len(warm) / t_warm
    rate = a.rate if a.rate > 0 else 0.85 * thr
    gaps = rng.exponential(1.0 / rate, size=a.requests)
    res, elapsed = run(list(range(a.requests)), gaps)
    lat = np.array([r[2]["latency"] for r in res]) * 1e3
    que = np.array([r[2]["queued"] for r in res]) * 1e3
    its = np.array([r[1]["iterations"] for r in res])
    trr = max(r[1]["true_rel_residual"] for r in res)
    first = min(r[2]["submit"] for r in res)
    last = max(r[2]["submit"] + r[2]["latency"] for r in res)
    st = srv.stats()
    srv.close()
    conv = np.array([r[1]["true_rel_residual"] <= 1e-8 for r in res])
    span = last - first
    # the plain FP64 per-system path on the first requests of the stream: one
    # request at a time (one slot), FP64 factor, exact single-launch solves,
    # the same equilibration, back to back; converged solves only
    plain = None
    if a.plain_requests > 0:
        sel = list(range(min(a.plain_requests, a.requests)))
        srv1 = ca.Server(slots=1, prec="f64", max_it=a.max_it or 3000,
                         solve_algorithm="single", equilibrate_sweeps=a.equilibrate_sweeps)
        used = sorted({int(picks[j]) for j in sel})
        reg = {}
        for i in used:
            k, p = _pattern(ca, int(sizes[i]), 1000 + i, a.max_block, order)
            reg[i] = srv1.add_pattern(p.row_ptr, p.col, p.val, p.piv_starts, p.block_starts)
        warm1 = srv1.submit(reg[used[0]], pats[used[0]][1][0], pats[used[0]][2])
        srv1.wait(warm1)
        t0 = time.perf_counter()
        ids1 = [srv1.submit(reg[int(picks[j])], pats[picks[j]][1][sets[j]], pats[picks[j]][2])
                for j in sel]
        out1 = [srv1.wait(r) for r in ids1]
        t1 = time.perf_counter() - t0
        srv1.close()
        c1 = sum(r[1]["true_rel_residual"] <= 1e-8 for r in out1)
        plain = {"requests": len(sel), "converged": int(c1), "solves_per_s": c1 / t1,
                 "sqmr_iterations_mean": float(np.mean([r[1]["iterations"] for r in out1])),
                 "max_true_rel_residual": max(r[1]["true_rel_residual"] for r in out1)}
        log(f"plain FP64 per-system path: {plain['solves_per_s']:.1f} solves/s over "
            f"{len(sel)} requests")
    return {
        "metric": "KKT solves/sec to rel. residual 1e-8 (FP32-factor vs FP64); p50/p99 solve latency",
        # only requests that reached the tolerance count as solves
        "value": float(conv.sum()) / span, "unit": "solves/s", "n_gpus": 1,
        "steps": len(res), "warmup": len(warm), "ms_per_step": span / len(res) * 1e3,
        "all_requests_per_s": len(res) / span,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if a.prec == "f64" else "f32-factor/f64-sqmr", "data": "synthetic",
        "config": {"workload": "serve", "patterns": a.patterns, "n_range": [a.n_lo, a.n_hi],
                   "m": "0.4n", "max_block": a.max_block, "slots": a.slots,
                   "requests": len(res), "offered_rate_per_s": rate,
                   "barrier_weights": [1e-4, 1e4], "tol": 1e-8, "max_it": a.max_it or 3000,
                   "solve_schedule": f"jacobi{sweeps}" if sweeps > 0 else exact,
                   "equilibrate_sweeps": a.equilibrate_sweeps, "graphs": bool(a.graphs),
                   "parallelism": "none"},
        "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
        "queue_p50_ms": float(np.percentile(que, 50)),
        "service_p50_ms": float(np.percentile(lat - que, 50)),
        "sqmr_iterations_mean": float(its.mean()), "max_true_rel_residual": trr,
        "converged_fraction": float(conv.mean()),
        "p50_ms_converged": float(np.percentile(lat[conv], 50)) if conv.any() else None,
        "p99_ms_converged": float(np.percentile(lat[conv], 99)) if conv.any() else None,
        "warmup_solves_per_s": thr, "setup_s": t_setup,
        "plain_f64_per_system": plain,
        "speedup_vs_plain_f64_per_system":
            None if not plain or plain["solves_per_s"] <= 0
            else float(conv.sum()) / span / plain["solves_per_s"],
        "device_bytes_peak": st["peak"], "device_budget": st["budget"],
        "instances": st["instances"],
        "roofline": None, "cpu_baseline": None,
    }


def bench_batch(a):
    """512 independent systems (n ~ U[1k, 40k], m = 0.4 n, seeds 0..511)
    solved as one batch (culip_amd.Batch): every system refactored
    concurrently on 16 streams, then one SQMR recurrence whose steps are
    single launches over all systems still iterating, with block-Jacobi
    triangular solves; the block size per size bucket ({8, 16, 32}) and the
    sweep count ({2, 3}) picked by the autotuner (culip_amd.batch). Beside it
    the plain FP64 path (exact single-launch solves), one system at a time,
    on the same systems. The value is the rate of a batch pass with the
    systems registered: refactor + solve of every system, values and
    right-hand sides from host memory and solutions back to it. Every
    system of a pass returns when the pass ends, so its latency is the pass
    time. The end-to-end rate also counts generating the systems,
    registering them (host symbolic analysis on 16 threads, upload) and the
    first pass, which allocates the pooled work space."""
    import culip_amd.batch as cb
    import concurrent.futures as cf
    rng = np.random.default_rng(a.seed)
    sizes = [int(v) for v in rng.integers(1000, 40001, size=a.systems)]
    t0 = time.perf_counter()
    tuned, sweeps, tune_ms = cb.autotune_batch(
        sizes, lambda i, B: cb.kkt_system(sizes[i], i, B), prec=a.prec, log=log)
    t_tune = time.perf_counter() - t0

    def make(i):
        k, p = cb.kkt_system(sizes[i], i, tuned[cb.bucket_of(sizes[i])])
        return p, k.rhs[p.perm].copy()

    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(16) as ex:
        systems = list(ex.map(make, range(a.systems)))
    t_gen = time.perf_counter() - t0
    log(f"{len(systems)} systems generated in {t_gen:.1f}s")
    res, secs, t_reg, phases = cb.solve_batch_engine(systems, a.prec, sweeps, log=log)
    el = secs[-1]
    # the same batch with the FP64 factor: batching alone, without the FP32
    # factor's halved factor bytes
    f64 = None
    if a.prec != "f64" and not a.no_plain:
        r64, s64, _, ph64 = cb.solve_batch_engine(systems, "f64", sweeps, log=log)
        f64 = {"value": len(r64) / s64[-1], "phases_s": {"refactors": ph64["factor_s"],
                                                         "sqmr": ph64["sqmr_s"]},
               "sqmr_iterations_mean": float(np.mean([r[1]["iterations"] for r in r64])),
               "sqmr_iterations_max": int(max(r[1]["iterations"] for r in r64)),
               "converged_fraction": float(np.mean([r[1]["true_rel_residual"] <= 1e-8
                                                    for r in r64]))}
    # the plain path: FP64 factor, exact single-launch solves (the reference
    # driver's SOLVE_BLOCK), one system at a time
    if a.no_plain:
        base, el_base = res, el
    else:
        base, secs_b, _ = cb.solve_batch(systems, lambda i: None, 1, "f64", log=log)
        el_base = secs_b[-1]
    trr = max(r[1]["true_rel_residual"] for r in res)
    conv = [r[1]["true_rel_residual"] <= 1e-8 for r in res]
    return {
        "metric": "KKT solves/sec to rel. residual 1e-8 (FP32-factor vs FP64); p50/p99 solve latency",
        "value": len(res) / el, "unit": "solves/s", "n_gpus": 1, "steps": len(res),
        "warmup": len(res), "ms_per_step": el / len(res) * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if a.prec == "f64" else "f32-factor/f64-sqmr", "data": "synthetic",
        "config": {"workload": "batch", "systems": a.systems, "n_range": [1000, 40000],
                   "m": "0.4n", "seeds": f"0..{a.systems - 1}",
                   "engine": "one batch: one launch per refactor step and per SQMR step over "
                             "all systems taking part",
                   "max_block_by_bucket": {f"{cb.BUCKETS[b][0]}-{cb.BUCKETS[b][1] - 1}":
                                           tuned[b] for b in tuned},
                   "solve_schedule": f"jacobi{sweeps}",
                   "tol": 1e-8, "parallelism": "none"},
        "p50_ms": el * 1e3, "p99_ms": el * 1e3,
        "latency_note": "every system of a pass returns when the pass ends",
        "max_true_rel_residual": trr,
        "converged_fraction": float(np.mean(conv)),
        "converged_solves_per_s": float(np.sum(conv)) / el,
        "sqmr_iterations_mean": float(np.mean([r[1]["iterations"] for r in res])),
        "sqmr_iterations_max": int(max(r[1]["iterations"] for r in res)),
        "plain_f64_per_system_solves_per_s": None if a.no_plain else len(base) / el_base,
        "speedup_vs_plain_f64_per_system":
            None if a.no_plain else (len(res) / el) / (len(base) / el_base),
        "plain_max_true_rel_residual":
            None if a.no_plain else max(r[1]["true_rel_residual"] for r in base),
        "phases_s": {"refactors": phases["factor_s"], "sqmr": phases["sqmr_s"]},
        "f64_factor_batch": f64,
        "device_bytes": phases["device_bytes"],
        "setup": {"autotune_s": t_tune, "generate_s": t_gen,
                  "register_s_symbolic_and_upload": t_reg,
                  "first_pass_s_with_allocation": secs[0]},
        "end_to_end_solves_per_s": len(res) / (t_gen + t_reg + secs[0]),
        "autotune_ms": tune_ms,
        "roofline": None, "cpu_baseline": None,
    }

def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {a.gpus}); "
              "independent KKT systems shard by system across GPUs with no exchange",
              file=sys.stderr)
        sys.exit(2)
    if a.slots <= 0:
        a.slots = 16 if a.workload == "batch" else 8
    # one hardware queue per slot stream, so concurrent solves are not
    # serialised behind each other (HIP maps streams onto this many queues;
    # read when the runtime starts); the default line also runs the batch
    # on 16 slots
    nq = a.slots if a.workload in ("serve", "batch") else 16
    os.environ.setdefault("GPU_MAX_HW_QUEUES", str(min(32, max(4, nq))))
    import torch
    import culip_amd as ca

    if a.workload in ("serve", "batch"):
        if a.prec is None:
            a.prec = "f32"
        if a.graphs < 0:
            a.graphs = 1 if a.workload == "serve" else 0
        out = bench_serve(a) if a.workload == "serve" else bench_batch(a)
        print(json.dumps(out))
        return
    if a.prec is None:
        a.prec = "f64"
    if a.graphs < 0:
        a.graphs = 1
    wl = WORKLOADS[a.workload]
    dev = torch.device("cuda:0")
    # one explicit stream for every launch (refactor graphs replay on
    # explicit streams only; the HIP events below record on it too)
    torch.cuda.set_stream(torch.cuda.Stream())
    kkt = ca.KKT(wl["n"], wl["m"], wl["o"], wl["c"], h_diag=wl["h_diag"], h_diag_log=wl["log"],
                 delta=wl["delta"], seed=a.seed, barrier=wl.get("barrier"))
    order = ca.ORDER_BLOCK_COLOR if a.order == "color" else ca.ORDER_NATURAL
    prep = ca.Prep(kkt.row_ptr, kkt.col, kkt.val, wl["n"], wl["B"], order)
    N = prep.row_ptr.size - 1
    Z = prep.col.size
    g = ca.BlockiLDLt(prep.row_ptr, prep.col, prep.val, prep.piv_starts, prep.block_starts,
                      a.prec)
    t_sym0 = time.perf_counter()
    g.symbolic(0, a.fill_factor)
    t_sym = time.perf_counter() - t_sym0
    sc = g.get_int("scalars")
    g.set_solve_algorithm(a.solve_algorithm, a.jacobi_sweeps)
    g.set_graphs(bool(a.graphs))
    A = ca.CSRMatrixSPMV(prep.row_ptr, prep.col, prep.val)
    solver = ca.SQMR(A, 1e-8, 1000, g, true_residual_period=20, max_restarts=10)
    dvals = torch.tensor(prep.val, device=dev)
    b_host = kkt.rhs[prep.perm]
    db = torch.tensor(b_host, device=dev)
    x = torch.zeros(N, dtype=torch.float64, device=dev)

    def step():
        g.refactor(dvals, a.threshold)
        solver.solve(db, x)

    if a.pmc_probe:
        # for hardware-counter runs: one step, then a plain device copy of
        # known size (the counters' bytes-per-KiB calibration)
        step()
        src = torch.ones(PROBE_COPY_DOUBLES, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"pmc_probe": a.workload, "copy_bytes_read": 8 * PROBE_COPY_DOUBLES,
                          "copy_bytes_written": 8 * PROBE_COPY_DOUBLES,
                          "sqmr_iterations": solver.result["iterations"]}))
        return

    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()
    lat, fac = [], []
    t0 = time.perf_counter()
    for _ in range(a.steps):
        s0 = time.perf_counter()
        g.refactor(dvals, a.threshold)
        s1 = time.perf_counter()
        solver.solve(db, x)
        s2 = time.perf_counter()
        lat.append(s2 - s0)
        fac.append(s1 - s0)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    res = solver.result
    elapsed = t1 - t0

    # dominant kernels, timed live with HIP events on the launch stream
    y = torch.empty_like(x)
    ms_spmv = event_time(lambda: A.multiply(x, y), 20)
    bytes_spmv = Z * 12 + (N + 1) * 4 + 2 * N * 8
    sv = 8 if a.prec == "f64" else 4
    ps = g.get_int("plan_sizes")
    # in place, as SQMR applies it: the launches of one U solve and nothing else
    y.copy_(db)
    ms_trsv = event_time(lambda: g.solve_right(y, y), 10)
    bytes_trsv = (int(ps[0]) + int(ps[1])) * (sv + 4) + 2 * N * 8 + 2 * (N + 1) * 4
    its = res["iterations"]
    per_it_spmv = ms_spmv
    peak = 8000.0
    rl_spmv = {"bound": "hbm", "kernel": "k_csr_spmv", "achieved": bytes_spmv / ms_spmv / 1e6,
               "peak": peak, "unit": "GB/s", "frac": bytes_spmv / ms_spmv / 1e6 / peak,
               "traffic": None, "bytes_per_launch": bytes_spmv, "ms_per_launch": ms_spmv}
    launches = int(sc[6])
    label = "k_trsv_level (one U solve, all levels)"
    if a.solve_algorithm == "jacobi":
        # sweeps after the first read every entry, the first only U_ii^-1
        launches = a.jacobi_sweeps
        first = int(ps[1]) * (sv + 4) + 2 * N * 8 + (N + 1) * 4
        bytes_trsv = first + (a.jacobi_sweeps - 1) * (bytes_trsv + N * 8)
        label = f"k_jacobi_sweep (one U solve, {a.jacobi_sweeps} sweeps)"
    elif a.solve_algorithm == "single":
        launches = 1
        label = "k_trsv_chain (one U solve)"
    rl_trsv = {"bound": "hbm", "kernel": label,
               "achieved": bytes_trsv / ms_trsv / 1e6, "peak": peak, "unit": "GB/s",
               "frac": bytes_trsv / ms_trsv / 1e6 / peak, "traffic": None,
               "bytes_per_launch": bytes_trsv, "ms_per_launch": ms_trsv,
               "launches": launches}
    # the U' solve that forms K ev itself, as SQMR applies it every
    # iteration (when the schedule folds the product; otherwise SQMR runs
    # the SpMV and the U' solve, both measured here)
    z = torch.empty_like(x)
    fused = g.left_product(A, x, z)
    rl_fused = None
    if fused:
        ms_fused = event_time(lambda: g.left_product(A, x, z), 10)
        bytes_fused = ((int(ps[2]) + int(ps[3])) * (sv + 4) + 2 * N * 8 + 2 * (N + 1) * 4
                       + Z * 12 + (N + 1) * 4 + N * 8)
        label_f = ("k_trsv_chain_product (one U' solve forming K ev)" if a.solve_algorithm == "single"
                   else f"k_jacobi_first_product + sweeps (one U' solve forming K ev)")
        rl_fused = {"bound": "hbm", "kernel": label_f,
                    "achieved": bytes_fused / ms_fused / 1e6, "peak": peak, "unit": "GB/s",
                    "frac": bytes_fused / ms_fused / 1e6 / peak, "traffic": None,
                    "bytes_per_launch": bytes_fused, "ms_per_launch": ms_fused,
                    "launches": 1 if a.solve_algorithm == "single" else a.jacobi_sweeps}
    tr = MEASURED_TRAFFIC.get((a.workload, a.prec, a.order, a.solve_algorithm))
    if tr is not None:
        rl_trsv["traffic"] = tr[label.split()[0]]
        rl_spmv["traffic"] = tr["k_csr_spmv"]
        if rl_fused is not None and "k_trsv_chain_product" in tr:
            rl_fused["traffic"] = tr["k_trsv_chain_product"]
        for r in (rl_trsv, rl_spmv, rl_fused):
            if r is not None:
                r["traffic_source"] = "profiles/kkt400k_hbm_traffic.md"
    # the dominant kernel: the one with the most time per SQMR iteration
    # (one U solve and one U' solve with the product per iteration; the
    # SpMV runs per iteration only where the product is not folded)
    per_it = [(ms_trsv, rl_trsv)]
    if rl_fused is not None:
        per_it.append((rl_fused["ms_per_launch"], rl_fused))
    else:
        per_it.append((per_it_spmv, rl_spmv))
    per_it.sort(key=lambda t: -t[0])
    dominant, secondary = per_it[0][1], per_it[1][1]

    def variant(prec, algorithm, sweeps):
        """The same step with another factor precision or solve schedule."""
        gv = ca.BlockiLDLt(prep.row_ptr, prep.col, prep.val, prep.piv_starts,
                           prep.block_starts, prec)
        gv.symbolic(0, a.fill_factor)
        gv.set_solve_algorithm(algorithm, sweeps)
        gv.set_graphs(bool(a.graphs))
        sv_ = ca.SQMR(A, 1e-8, 1000, gv, true_residual_period=20, max_restarts=10)
        xv = torch.zeros_like(x)
        for _ in range(a.warmup):
            gv.refactor(dvals, a.threshold)
            sv_.solve(db, xv)
        torch.cuda.synchronize()
        tv = time.perf_counter()
        for _ in range(a.steps):
            gv.refactor(dvals, a.threshold)
            sv_.solve(db, xv)
        torch.cuda.synchronize()
        tv = time.perf_counter() - tv
        rv = sv_.result
        return {"value": a.steps / tv, "ms_per_step": tv / a.steps * 1e3, "prec": prec,
                "solve_algorithm": algorithm, "jacobi_sweeps": sweeps,
                "sqmr_iterations": rv["iterations"], "refinement_restarts": rv["restarts"],
                "true_rel_residual": rv["true_rel_residual"],
                "rel_diff_vs_fp64_solution": float((xv - x).norm() / x.norm())}

    # FP32 factor (FP64 SQMR, FP64 true residual); block-Jacobi solves, with
    # the FP64 and the FP32 factor (the sweeps read the factor's entries in
    # its storage precision, so that is where the FP32 factor can pay)
    other = jac = jac32 = None
    if a.prec == "f64" and not a.no_variants:
        other = variant("f32", a.solve_algorithm, a.jacobi_sweeps)
        if a.solve_algorithm in ("single", "levels"):
            log("variant: jacobi solves, 2 sweeps")
            jac = variant("f64", "jacobi", 2)
            jac32 = variant("f32", "jacobi", 2)

    prec_var = None
    if not a.no_variants and not a.no_cpu_baseline:
        log("precision variant: barrier-conditioned n=2,000, FP32 factor vs the oracle")
        prec_var = precision_variant(order, a.seed, a.fill_factor, a.threshold)

    # the batching configuration (512 systems, FP32 factor, autotuned), in
    # the same line so that every run of the default bench measures it
    batch = None
    if not a.no_variants:
        log(f"batch: {a.systems} systems")
        ab = argparse.Namespace(**vars(a))
        ab.workload, ab.slots, ab.prec, ab.graphs = "batch", 16, "f32", 0
        bo = bench_batch(ab)
        batch = {k: bo[k] for k in ("value", "unit", "converged_fraction", "max_true_rel_residual",
                                    "sqmr_iterations_mean", "p50_ms", "p99_ms",
                                    "plain_f64_per_system_solves_per_s",
                                    "speedup_vs_plain_f64_per_system",
                                    "end_to_end_solves_per_s", "phases_s", "f64_factor_batch",
                                    "setup")}
        batch["config"] = bo["config"]

    cpu = None
    if not a.no_cpu_baseline:
        log("cpu baseline: oracle on the host")
        try:
            cpu = cpu_baseline(wl, order, a.seed, a.fill_factor, a.threshold)
        except Exception as e:  # the oracle is optional on the box
            cpu = {"error": str(e)}

    lat_ms = np.array(lat) * 1e3
    out = {
        "metric": "KKT solves/sec to rel. residual 1e-8 (FP32-factor vs FP64); p50/p99 solve latency",
        "value": a.steps / elapsed,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": elapsed / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if a.prec == "f64" else "f32-factor/f64-sqmr",
        "data": "synthetic",
        "config": {"workload": a.workload, "n": wl["n"], "m": wl["m"], "N": N, "nnz": Z,
                   "max_block": wl["B"], "fill_level": 0, "fill_factor": a.fill_factor,
                   "drop_tol": a.threshold, "ordering": a.order, "levels": int(sc[6]),
                   "coarse_blocks": int(sc[5]), "tol": 1e-8, "parallelism": "none",
                   "solve_algorithm": a.solve_algorithm, "jacobi_sweeps": a.jacobi_sweeps,
                   "refactor_graphs": bool(a.graphs)},
        "p50_ms": float(np.percentile(lat_ms, 50)),
        "p99_ms": float(np.percentile(lat_ms, 99)),
        "factor_ms_mean": float(np.mean(fac) * 1e3),
        "sqmr_iterations": its,
        "refinement_restarts": res["restarts"],
        "true_rel_residual": res["true_rel_residual"],
        "symbolic_s_once_per_pattern": t_sym,
        "roofline": dominant,
        "roofline_secondary": secondary,
        "roofline_spmv": rl_spmv,
        # no MFMA instruction runs: at level-0 fill on this generator 2.74M
        # of the 2.77M off-diagonal blocks hold one entry and take 1.9
        # Schur pairs each, so every Schur update is a scalar multiply-add
        "mfma_util": 0.0,
        "mfma_util_note": "no MFMA path: Schur updates at level-0 fill are 1-entry blocks "
                          "(2.74M of 2.77M off-diagonal blocks, 1.9 pairs each), one scalar "
                          "FMA per update; nothing to contract",
        "cpu_baseline": cpu,
    }
    if other is not None:
        out["f32_factor"] = other
    if jac is not None:
        out["jacobi_solves"] = jac
    if jac32 is not None:
        out["f32_factor_jacobi_solves"] = jac32
    if prec_var is not None:
        out["precision_variant"] = prec_var
    if batch is not None:
        out["batch"] = batch
    print(json.dumps(out))


if __name__ == "__main__":
    main()
