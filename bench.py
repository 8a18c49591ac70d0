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
    ap.add_argument("--workload", default="kkt-400k", choices=sorted(WORKLOADS))
    ap.add_argument("--prec", default="f64", choices=["f64", "f32"])
    ap.add_argument("--order", default="color", choices=["color", "natural"])
    ap.add_argument("--fill-factor", type=float, default=1.0)
    ap.add_argument("--threshold", type=float, default=1e-4)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    return ap.parse_args()


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


def cpu_baseline(wl, order, seed, fill_factor, threshold):
    """The oracle (C restatement of the reference algorithm, one core): a full
    solve (symbolic, factor, SQMR with the reference's stopping test) of the
    same generator at 1/scale of the workload's n and m, timed and scaled
    linearly by nnz(K) to the workload (kkt-400k: scale 10, ~20 s of CPU)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import culip_amd as ca
    import oracle_lib as ol
    scale = 10 if wl["n"] >= 100000 else 1
    n, m = wl["n"] // scale, wl["m"] // scale
    kkt = ca.KKT(n, m, wl["o"], wl["c"], h_diag=wl["h_diag"], h_diag_log=wl["log"],
                 delta=wl["delta"], seed=seed, barrier=wl.get("barrier"))
    prep = ca.Prep(kkt.row_ptr, kkt.col, kkt.val, n, wl["B"], order)
    b = kkt.rhs[prep.perm]
    o = ol.OracleBildlt(prep.row_ptr, prep.col, prep.val, prep.piv_starts, prep.block_starts,
                        "d")
    t0 = time.perf_counter()
    o.symbolic(0, fill_factor)
    o.factor(threshold)
    t1 = time.perf_counter()
    x, r = ol.sqmr_bildlt(prep.row_ptr, prep.col, prep.val, o, b, 1e-8, 1000)
    t2 = time.perf_counter()
    nnz_full = wl["n"] * (1 + wl["o"]) + 2 * wl["c"] * wl["n"] + wl["m"]
    ratio = nnz_full / prep.col.size
    t_solve = (t2 - t0) * ratio
    return {"value": 1.0 / t_solve, "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": f"oracle full solve of the same generator at n={n}, m={m} (seed {seed}): "
                      f"symbolic+factor {t1 - t0:.2f}s, SQMR {r.iterations} iterations "
                      f"{t2 - t1:.2f}s; scaled by nnz ratio {ratio:.2f} to the workload"}


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {a.gpus}); "
              "independent KKT systems shard by system across GPUs with no exchange",
              file=sys.stderr)
        sys.exit(2)
    import torch
    import culip_amd as ca

    wl = WORKLOADS[a.workload]
    dev = torch.device("cuda:0")
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
    A = ca.CSRMatrixSPMV(prep.row_ptr, prep.col, prep.val)
    solver = ca.SQMR(A, 1e-8, 1000, g, true_residual_period=20, max_restarts=10)
    dvals = torch.tensor(prep.val, device=dev)
    b_host = kkt.rhs[prep.perm]
    db = torch.tensor(b_host, device=dev)
    x = torch.zeros(N, dtype=torch.float64, device=dev)

    def step():
        g.refactor(dvals, a.threshold)
        solver.solve(db, x)

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
    ms_trsv = event_time(lambda: g.solve_right(db, y), 10)
    bytes_trsv = (int(ps[0]) + int(ps[1])) * (sv + 4) + 2 * N * 8 + 2 * (N + 1) * 4
    its = res["iterations"]
    per_it_spmv = ms_spmv
    per_it_trsv = 2 * ms_trsv
    peak = 8000.0
    rl_spmv = {"bound": "hbm", "kernel": "k_csr_spmv", "achieved": bytes_spmv / ms_spmv / 1e6,
               "peak": peak, "unit": "GB/s", "frac": bytes_spmv / ms_spmv / 1e6 / peak,
               "traffic": None, "bytes_per_launch": bytes_spmv, "ms_per_launch": ms_spmv}
    rl_trsv = {"bound": "hbm", "kernel": "k_trsv_level (one U solve, all levels)",
               "achieved": bytes_trsv / ms_trsv / 1e6, "peak": peak, "unit": "GB/s",
               "frac": bytes_trsv / ms_trsv / 1e6 / peak, "traffic": None,
               "bytes_per_launch": bytes_trsv, "ms_per_launch": ms_trsv,
               "launches": int(sc[6])}
    dominant, secondary = (rl_trsv, rl_spmv) if per_it_trsv >= per_it_spmv else (rl_spmv, rl_trsv)

    cpu = None
    if not a.no_cpu_baseline:
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
                   "coarse_blocks": int(sc[5]), "tol": 1e-8, "parallelism": "none"},
        "p50_ms": float(np.percentile(lat_ms, 50)),
        "p99_ms": float(np.percentile(lat_ms, 99)),
        "factor_ms_mean": float(np.mean(fac) * 1e3),
        "sqmr_iterations": its,
        "refinement_restarts": res["restarts"],
        "true_rel_residual": res["true_rel_residual"],
        "symbolic_s_once_per_pattern": t_sym,
        "roofline": dominant,
        "roofline_secondary": secondary,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
