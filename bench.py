#!/usr/bin/env python3
"""Headline benchmark of the multi-evaluation hot path on one MI355X.

Workload (the streaming configuration the metric is quoted on): ModLinear
grid at d=10, regular level 4 refined by +200 surplus refinements per step of
the hierarchised f(x) = exp(-sum_j (x_j - 0.5)^2 / 0.05) until >= 100,000
points; 1,000,000 data points from a 5-component Gaussian mixture (sigma
0.08, seed 3). A step is one mult (B alpha) plus one multTranspose (B^T y)
over all M x N pairs, with inputs resident in HBM.

`value` is the FP32-with-FP64-accumulation form's pair-evals/s (2*M*N per
step); the plain FP64 kernel is timed beside it ("fp64", "speedup_vs_fp64").
Both forms' relative errors against the oracle are reported on a sample.
roofline: the dominant FP32 kernel's executed flops (counted in the run by
the counting instantiation of the same kernel) over its HIP-event time,
against the FP32 vector peak; the 6d-per-pair figure over all M*N pairs is
reported beside it as achieved_6d_convention.
cpu_baseline: the reference's StreamingDefault ModLinear operator
(OperationMultiEvalModMaskStreaming, compiled from the reference sources into
oracle/_ref) on the host cores, FP64, on a bounded data sample (warm-up, then
the median of 5).
served: the served-loop workload (64 resident models, Poisson requests) for
a short stream at 2000 and at 8000 req/s, the FP32 continuous-batching server
against one launch per request in plain FP64: evals/s, p50/p99 latency.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu] [--no-serve]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3}  # MI355X published vector peaks


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-check", action="store_true", help="skip the accuracy sample")
    ap.add_argument("--cpu-seconds", type=float, default=4.0,
                    help="target seconds per CPU mult+multTranspose run")
    ap.add_argument("--serve", action="store_true",
                    help="run the served-loop workload (64 models, Poisson requests) instead")
    ap.add_argument("--serve-seconds", type=float, default=4.0)
    ap.add_argument("--no-autotune", action="store_true",
                    help="serve with the untuned default tile shapes")
    ap.add_argument("--serve-rate", type=float, default=2000.0,
                    help="Poisson arrival rate of the served workload (requests/s)")
    ap.add_argument("--serve-baseline-kernel-choice", action="store_true",
                    help="also run the one-launch baseline with the served loop's per-model "
                         "kernel choice")
    ap.add_argument("--serve-no-baseline", action="store_true",
                    help="serve only; skip the one-launch-per-request FP64 baseline")
    ap.add_argument("--serve-autotune-cache", default=None,
                    help="file of tuned tile shapes, read and updated by both servers")
    ap.add_argument("--serve-operator-sample", type=int, default=1000,
                    help="requests timed with an operator built per request (the reference's "
                         "predict pattern); 0 skips it")
    ap.add_argument("--solve", action="store_true",
                    help="run the regression-solve workload (FP64 CG vs mixed precision) instead")
    ap.add_argument("--inner-epsilon", type=float, default=1e-5)
    ap.add_argument("--large", action="store_true",
                    help="run the large streaming workload (Linear d=20, N>=1e6, M=1e7) instead")
    ap.add_argument("--grid-target-large", type=int, default=1000000)
    ap.add_argument("--data-points-large", type=int, default=10000000)
    ap.add_argument("--copy-reference-mib", type=int, default=0,
                    help="first copy this many MiB device to device three times (a kernel "
                         "of known traffic for scaling hardware byte counters)")
    ap.add_argument("--no-serve", action="store_true",
                    help="skip the served-loop leg of the default run")
    ap.add_argument("--serve-leg-seconds", type=float, default=1.5,
                    help="seconds of each request stream (2000 and 16000 req/s) in the default run")
    ap.add_argument("--grid-target", type=int, default=100000)
    ap.add_argument("--data-points", type=int, default=1000000)
    return ap.parse_args()


def timed_steps(op, prec, alpha, y, r, a, steps, warmup, torch):
    for _ in range(warmup):
        op.mult_device(alpha.data_ptr(), r.data_ptr(), prec)
        op.mult_transpose_device(y.data_ptr(), a.data_ptr(), prec)
    op.synchronize()
    op.kernel_timing_enable(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        op.mult_device(alpha.data_ptr(), r.data_ptr(), prec)
        op.mult_transpose_device(y.data_ptr(), a.data_ptr(), prec)
    op.synchronize()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    nm, msm, nt, mst = op.kernel_timing()
    op.kernel_timing_enable(False)
    return dt / steps, msm / nm, mst / nt


def accuracy(basis, X, lv, ix, alpha, y, results):
    """relative errors of each form on a sample, against the oracle"""
    import numpy as np

    import oracle
    rows = np.arange(0, X.shape[0], max(1, X.shape[0] // 2000))[:2000]
    cols = np.arange(0, lv.shape[0], max(1, lv.shape[0] // 100))[:100]
    r0 = oracle.mult(basis, X[rows], lv, ix, alpha)
    a0 = oracle.mult_transpose(basis, X, lv[cols], ix[cols], y)
    out = {}
    for name, (r, a) in results.items():
        out[name] = {
            "mult_rel_l2": float(np.linalg.norm(r[rows] - r0) / np.linalg.norm(r0)),
            "multT_rel_l2": float(np.linalg.norm(a[cols] - a0) / np.linalg.norm(a0)),
            "mult_max_abs_over_max": float(np.max(np.abs(r[rows] - r0)) / np.max(np.abs(r0))),
            "multT_max_abs_over_max": float(np.max(np.abs(a[cols] - a0)) / np.max(np.abs(a0))),
        }
    out["sample"] = "%d data rows (mult), %d grid points (multTranspose); oracle = long-double definition" % (
        len(rows), len(cols))
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(basis, X, grid, lv, ix, alpha, y, seconds):
    """The reference's StreamingDefault operator (or the C restatement when the
    reference build is absent) on the host cores: one warm-up run, then the
    median of 5 timed mult+multTranspose runs on a data sample sized to about
    `seconds` per run."""
    import numpy as np

    import oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or oracle.num_threads()
    N = lv.shape[0]
    kind = "reference" if oracle.ref() is not None else "port"
    # the reference operator reads the same points, inserted in the same order
    rgrid = oracle.RefGrid.from_arrays(basis, lv, ix) if kind == "reference" else None

    def make(m):
        if kind == "reference":
            return oracle.RefOp(rgrid, oracle.KIND_STREAMING, X[:m])
        return oracle.Port(basis, X[:m], lv, ix)

    def run(op, m):
        t = time.perf_counter()
        op.mult(alpha)
        op.mult_transpose(y[:m])
        return time.perf_counter() - t

    m0 = 20000
    op = make(m0)
    run(op, m0)
    t0 = run(op, m0)
    m = int(min(X.shape[0], max(m0, m0 * seconds / max(t0, 1e-3))))
    op = make(m)
    run(op, m)  # warm-up
    times = sorted(run(op, m) for _ in range(5))
    med = times[2]
    return {
        "value": 2.0 * m * N / med,
        "unit": "pair-evals/s",
        "cores": threads,
        "kind": kind,
        "cpu": cpu_model(),
        "run_seconds": times,
        "sample": ("%s StreamingDefault ModLinear operator (OperationMultiEvalModMaskStreaming%s), "
                   "FP64, full %d-point grid x first %d data points, mult+multTranspose, one "
                   "warm-up then the median of 5"
                   % ("reference" if kind == "reference" else "restated",
                      ", %s build" % oracle.ref_variant() if kind == "reference" else "", N, m)),
    }


def drive_server(srv, ids, c, warm=32, clients=4):
    """Replay the request stream in real time against one server, submitted
    by `clients` threads (request k by thread k mod clients, each at its
    arrival time; the C-ABI submit is thread-safe and converts the queries on
    the calling thread); returns per-request latency (arrival -> result),
    completion times and outputs."""
    import queue
    import threading

    import numpy as np
    dims = [m["dim"] for m in c["models"]]
    for k in range(warm):  # warm-up outside the timed trace
        m = int(c["model"][k])
        X = c["pools"][dims[m]][:64]
        rid, q = srv.submit(ids[m], X)
        srv.poll(rid, q)
    n = len(c["arrival"])
    lat = np.full(n, np.nan)
    status = np.zeros(n, dtype=np.int64)
    outs = {}
    keep = set(range(0, n, max(1, n // 64)))
    pend = queue.Queue()
    t0 = time.perf_counter() + 0.05

    def submitter(j):
        for k in range(j, n, clients):
            ta = t0 + float(c["arrival"][k])
            now = time.perf_counter()
            if ta > now:
                time.sleep(ta - now)
            m, q, off = int(c["model"][k]), int(c["q"][k]), int(c["offset"][k])
            X = c["pools"][dims[m]][off:off + q]
            lag = time.perf_counter() - ta
            rid, _ = srv.submit(ids[m], X)
            pend.put((k, rid, q, lag))
        pend.put(None)

    ths = [threading.Thread(target=submitter, args=(j,)) for j in range(clients)]
    for th in ths:
        th.start()
    ended = 0
    lags = np.zeros(n)
    while ended < clients:
        item = pend.get()
        if item is None:
            ended += 1
            continue
        k, rid, q, lag = item
        st, out, l = srv.poll(rid, q, wait_s=-1.0)
        status[k] = st
        lat[k] = lag + l
        lags[k] = lag
        if k in keep and out is not None:
            outs[k] = out
    for th in ths:
        th.join()
    done = status == 2
    t_complete = c["arrival"] + lat
    span = float(np.nanmax(t_complete[done]) - c["arrival"][0])
    evals = float(c["q"][done].sum())
    return dict(evals_per_s=evals / span, requests_done=int(done.sum()), requests=n,
                p50_ms=float(np.percentile(lat[done], 50) * 1e3),
                p99_ms=float(np.percentile(lat[done], 99) * 1e3),
                mean_ms=float(np.mean(lat[done]) * 1e3), span_s=span,
                submit_lag_ms={"mean": float(np.mean(lags) * 1e3),
                               "max": float(np.max(lags) * 1e3)},
                clients=clients, stats=srv.stats()), outs


def operator_per_request(c, n):
    """The reference's prediction pattern (LearnerLeastSquaresIdentity::predict,
    LearnerLeastSquaresIdentity.cpp:61-68): per request, an FP64 operator is
    built on the query points, the model's grid is prepared and mult runs;
    host arrays in and out. Closed loop over the first n requests of the
    stream (no arrival times): evals/s = queries / wall time."""
    from sgpp_amd import capi
    dims = [m["dim"] for m in c["models"]]
    lvix = [(m["level"], m["index"]) for m in c["models"]]
    n = min(n, len(c["q"]))
    t0 = time.perf_counter()
    evals = 0
    for k in range(n):
        m, q, off = int(c["model"][k]), int(c["q"][k]), int(c["offset"][k])
        X = c["pools"][dims[m]][off:off + q]
        op = capi.MultiEval(c["models"][m]["basis"], X, precision=capi.PREC_FP64)
        op.prepare(*lvix[m])
        op.mult(c["models"][m]["alpha"])
        del op
        evals += q
        if (k + 1) % 250 == 0:
            print("[serve] operator per request: %d done, %.1f s" % (k + 1, time.perf_counter() - t0),
                  file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    return {"requests": n, "queries": evals, "wall_s": wall, "evals_per_s": evals / wall,
            "mean_ms_per_request": wall / n * 1e3}


def thin_stream(c, frac, seed=12):
    """Poisson thinning: each request of the stream kept with probability frac
    (a seeded draw), which leaves a Poisson stream at frac times the rate."""
    import numpy as np

    from sgpp_amd import workloads
    keep = workloads.uniform(workloads.stream(5, seed), len(c["arrival"])) < frac
    out = dict(c)
    for k in ("arrival", "model", "q", "offset"):
        out[k] = np.asarray(c[k])[keep]
    out["rate"] = c["rate"] * frac
    return out


def serve_leg(rates, seconds, baseline=True, autotune=True, cache=None, baseline_auto=False):
    """The served workload (64 resident models, Poisson requests) at each rate
    of `rates` for `seconds`: one stream is drawn at the highest rate and the
    lower rates are thinned from it. The FP32 server (continuous batching;
    with the tuner, each model on the faster of the streaming and the
    subspace kernel) and the baseline (one launch per request, the plain FP64
    streaming kernel, same resident models) each admit the models once and
    replay every stream in real time. baseline_auto adds the one-launch
    baseline with the same per-model kernel choice as the served loop.
    Returns (per-rate results, the stream at the highest rate, outputs)."""
    from sgpp_amd import capi, workloads
    rates = sorted(rates)
    t = time.perf_counter()
    c_hi = workloads.config_serving(rate=rates[-1], duration=seconds)
    gen_s = time.perf_counter() - t
    streams = {r: (c_hi if r == rates[-1] else thin_stream(c_hi, r / rates[-1])) for r in rates}
    configs = [("served", dict(precision=capi.PREC_FP32, max_inflight=2))]
    if baseline:
        configs.append(("one_launch_per_request_fp64",
                        dict(precision=capi.PREC_FP64, max_inflight=1, max_requests_per_batch=1,
                             use_graph=False, algorithm=capi.SERVE_STREAMING)))
    if baseline_auto:
        configs.append(("one_launch_per_request_fp64_kernel_choice",
                        dict(precision=capi.PREC_FP64, max_inflight=1, max_requests_per_batch=1,
                             use_graph=False, algorithm=capi.SERVE_AUTO)))
    algorithms = {}
    runs = {r: {} for r in rates}
    outputs = {r: {} for r in rates}
    for name, kw in configs:
        kw["autotune"] = autotune
        kw["autotune_cache"] = cache
        t = time.perf_counter()
        srv = capi.Server(**kw)
        ids = [srv.admit_model(m["basis"], m["level"], m["index"], m["alpha"])
               for m in c_hi["models"]]
        admit_s = time.perf_counter() - t
        n_sub = sum(srv.model_algorithm(i) == capi.SERVE_SUBSPACE for i in ids)
        algorithms[name] = {"subspace": n_sub, "streaming": len(ids) - n_sub}
        print("[serve] %s: %d models admitted in %.1f s (%d on the subspace kernel)"
              % (name, len(ids), admit_s, n_sub), file=sys.stderr, flush=True)
        for r in rates:
            res, outs = drive_server(srv, ids, streams[r])
            res["admit_and_tune_s"] = admit_s
            runs[r][name] = res
            outputs[r][name] = outs
            print("[serve] %s at %.0f req/s: %.3g evals/s, p99 %.1f ms" %
                  (name, r, res["evals_per_s"], res["p99_ms"]), file=sys.stderr, flush=True)
        if name == "served":
            runs["tuned_shapes"] = {
                "%s_d%d" % ("modlinear" if b else "linear", d): srv.tuned_shape(b, d)
                for b, d in sorted({(m["basis"], m["dim"]) for m in c_hi["models"]})
                if srv.tuned_shape(b, d) is not None}
        srv.close()
    import numpy as np
    N = np.array([m["level"].shape[0] for m in c_hi["models"]])
    D = np.array([m["dim"] for m in c_hi["models"]])
    out = {"generation_s": gen_s, "seconds": seconds, "tuned_shapes": runs.pop("tuned_shapes", None),
           "models_per_kernel": algorithms, "rates": {}}
    for r in rates:
        c = streams[r]
        flops = float(sum(6 * D[m] * N[m] * q for m, q in zip(c["model"], c["q"])))
        sv = runs[r]["served"]
        sv["achieved_tflops_6d"] = (flops * sv["requests_done"] / len(c["q"]) / sv["span_s"]
                                    / 1e12)
        base = runs[r].get("one_launch_per_request_fp64")
        ref = outputs[r].get("one_launch_per_request_fp64", {})
        errs = [np.linalg.norm(outputs[r]["served"][k] - ref[k]) /
                max(np.linalg.norm(ref[k]), 1e-300) for k in outputs[r]["served"] if k in ref]
        out["rates"]["%d" % int(r)] = {
            "requests": len(c["q"]), "mean_queries": float(c["q"].mean()),
            "offered_evals_per_s": float(c["q"].sum() / seconds),
            "offered_tflops_6d": flops / seconds / 1e12,
            "served": sv, "baseline": base,
            "baseline_kernel_choice": runs[r].get("one_launch_per_request_fp64_kernel_choice"),
            "evals_ratio_vs_baseline": sv["evals_per_s"] / base["evals_per_s"] if base else None,
            "p99_ratio_vs_baseline": sv["p99_ms"] / base["p99_ms"] if base else None,
            "max_rel_err_vs_fp64": float(max(errs)) if errs else None,
        }
    return out, c_hi


def serve_bench(args):
    rates = [args.serve_rate]
    leg, c = serve_leg(rates, args.serve_seconds, baseline=not args.serve_no_baseline,
                       autotune=not args.no_autotune, cache=args.serve_autotune_cache,
                       baseline_auto=args.serve_baseline_kernel_choice)
    r = leg["rates"]["%d" % int(args.serve_rate)]
    N = [m["level"].shape[0] for m in c["models"]]
    line = {
        "metric": "served evals/s + p50/p99 latency",
        "value": r["served"]["evals_per_s"],
        "unit": "evals/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "p50_ms": r["served"]["p50_ms"],
        "p99_ms": r["served"]["p99_ms"],
        "dtype": "fp32 products with fp64 accumulation (streaming kernel), fp64 (subspace kernel)",
        "data": "synthetic, seeded: 64 models, Poisson %.0f req/s for %.1f s" % (args.serve_rate,
                                                                              args.serve_seconds),
        "config": {"workload": "served_64_models_poisson_%dreqs" % int(args.serve_rate),
                   "models": len(c["models"]), "grid_points_total": int(sum(N)),
                   "requests": r["requests"], "mean_queries": r["mean_queries"],
                   "offered_evals_per_s": r["offered_evals_per_s"],
                   "offered_tflops_6d": r["offered_tflops_6d"]},
        "served": r["served"],
        "baseline": r["baseline"],
        "baseline_kernel_choice": r["baseline_kernel_choice"],
        "models_per_kernel": leg["models_per_kernel"],
        "evals_ratio_vs_baseline": r["evals_ratio_vs_baseline"],
        "p99_ratio_vs_baseline": r["p99_ratio_vs_baseline"],
        "max_rel_err_vs_fp64": r["max_rel_err_vs_fp64"],
        "tuned_shapes": leg["tuned_shapes"],
        "setup_s": leg["generation_s"],
    }
    if args.serve_operator_sample > 0:
        print("[serve] operator per request: %d requests" % args.serve_operator_sample,
              file=sys.stderr, flush=True)
        opr = operator_per_request(c, args.serve_operator_sample)
        line["operator_per_request_fp64"] = opr
        line["evals_ratio_vs_operator_per_request"] = (r["served"]["evals_per_s"] /
                                                       opr["evals_per_s"])
    print(json.dumps(line))


def solve_bench(args):
    """Regression solve (Linear d=8 level 6, 500,000 points, lambda 1e-4) to a
    1e-10 relative residual: plain FP64 CG against FP32 products with FP64 CG
    vectors and FP64 iterative refinement, alpha checked against the reference's
    FP64 CPU solve (tests/golden/regression_alpha.npz) when present."""
    import numpy as np
    import torch

    from sgpp_amd import capi, workloads
    copy_reference(args.copy_reference_mib)
    c = workloads.config_regression()
    lv, ix = c["grid"].level_index()
    X, y, lam = c["X"], c["y"], c["lam"]
    M, N, d = X.shape[0], lv.shape[0], lv.shape[1]
    op = capi.MultiEval(capi.BASIS_LINEAR, X, precision=capi.PREC_FP32, keep_fp64=True)
    op.prepare(lv, ix)
    dev = torch.device("cuda:0")
    d_a = torch.tensor(workloads.normal(9, N), device=dev)
    d_y = torch.tensor(y, device=dev)
    d_r = torch.empty(M, dtype=torch.float64, device=dev)
    d_b = torch.empty(N, dtype=torch.float64, device=dev)
    kern = {}
    for name, prec in (("fp64", capi.PREC_FP64), ("fp32", capi.PREC_FP32)):
        step_s, mult_ms, multT_ms = timed_steps(op, prec, d_a, d_y, d_r, d_b, 5, 1, torch)
        kern[name] = {"pair_evals_per_s": 2.0 * M * N / step_s, "mult_ms": mult_ms,
                      "multT_ms": multT_ms}
    ref = None
    rpath = os.path.join(ROOT, "tests", "golden", "regression_alpha.npz")
    if os.path.exists(rpath):
        ref = np.load(rpath)
    runs = {}
    for name, kw in (("fp64_cg", dict(mode=capi.SOLVE_FP64)),
                     ("mixed_refinement", dict(mode=capi.SOLVE_MIXED,
                                               inner_epsilon=args.inner_epsilon,
                                               max_refinements=40))):
        alpha, info = op.solve(y, lam, max_iterations=10000, epsilon=1e-10, **kw)
        pairs = 2.0 * M * N * (info["fp64_matvecs"] + info["fp32_matvecs"])
        info["pair_evals_per_s"] = pairs / info["seconds"]
        if ref is not None:
            info["alpha_rel_err_vs_cpu_fp64"] = float(np.linalg.norm(alpha - ref["alpha"]) /
                                                      np.linalg.norm(ref["alpha"]))
        runs[name] = info
    # the subspace-indexed operator (SubspaceLinear semantics, FP64): kernels
    # and the same FP64 CG solve
    sop = capi.MultiEval(capi.BASIS_LINEAR, X, algorithm=capi.ALGO_SUBSPACE).prepare(lv, ix)
    step_s, mult_ms, multT_ms = timed_steps(sop, capi.PREC_FP64, d_a, d_y, d_r, d_b, 5, 1, torch)
    kern["subspace_fp64"] = {"pair_evals_per_s": 2.0 * M * N / step_s, "mult_ms": mult_ms,
                             "multT_ms": multT_ms, "executed_work": sop.count_work(capi.PREC_FP64)}
    alpha_s, info_s = sop.solve(y, lam, max_iterations=10000, epsilon=1e-10, mode=capi.SOLVE_FP64)
    if ref is not None:
        info_s["alpha_rel_err_vs_cpu_fp64"] = float(np.linalg.norm(alpha_s - ref["alpha"]) /
                                                    np.linalg.norm(ref["alpha"]))
    runs["fp64_cg_subspace"] = info_s
    del sop
    line = {
        "metric": "regression solve time to a 1e-10 relative residual, FP32+FP64 refinement vs FP64",
        "value": runs["fp64_cg"]["seconds"] / runs["mixed_refinement"]["seconds"],
        "unit": "x speedup (FP64 CG time / mixed time)",
        "n_gpus": 1,
        "higher_is_better": True,
        "config": {"workload": c["name"], "grid_points": N, "data_points": M, "dim": d,
                   "lambda": lam, "epsilon": 1e-10, "inner_epsilon": args.inner_epsilon},
        "kernels": kern,
        "kernel_speedup_fp32_vs_fp64": kern["fp32"]["pair_evals_per_s"] / kern["fp64"]["pair_evals_per_s"],
        "fp64_cg": runs["fp64_cg"],
        "mixed_refinement": runs["mixed_refinement"],
        "fp64_cg_subspace_operator": runs["fp64_cg_subspace"],
        "cpu_reference": ({"iterations": int(ref["iterations"]), "seconds": float(ref["seconds"]),
                           "threads": int(ref["threads"]),
                           "solver": "reference ConjugateGradients on OperationMultipleEvalSubspaceCombined",
                           "where": "build container, when the fixture was written"}
                          if ref is not None else None),
    }
    if not args.no_cpu:
        line["cpu_reference_this_host"] = cpu_solve_slice(c, lv, ix,
                                                          int(ref["iterations"]) if ref is not None
                                                          else None)
    print(json.dumps(line))


def cpu_solve_slice(c, lv, ix, iterations):
    """The reference's FP64 CPU solve (ConjugateGradients on its SubspaceLinear
    operator, compiled from its sources into oracle/_ref) timed on this host:
    one CG iteration and four (each including b = B^T y and the start
    residual), whose difference gives the time per iteration; the full solve is
    projected from the fixture's iteration count."""
    import oracle
    if oracle.ref() is None:
        return None
    g = oracle.RefGrid.from_arrays(0, lv, ix)
    t = time.perf_counter()
    op = oracle.RefOp(g, oracle.KIND_SUBSPACE, c["X"])
    setup = time.perf_counter() - t
    times = {}
    for k in (1, 4):
        t = time.perf_counter()
        op.cg_solve(c["y"], c["lam"], k, 1e-10)
        times[k] = time.perf_counter() - t
    per_it = (times[4] - times[1]) / 3.0
    out = {"threads": oracle.num_threads(), "cpu": cpu_model(), "operator_setup_s": setup,
           "solve_1_iteration_s": times[1], "solve_4_iterations_s": times[4],
           "seconds_per_iteration": per_it}
    if iterations:
        out["projected_solve_s"] = times[1] + (iterations - 1) * per_it
        out["projected_iterations"] = iterations
    return out


def large_bench(args):
    """Large streaming case: Linear d=20 grid refined to >= 1,000,000 points,
    10,000,000 FP32-representable mixture points, 5 CG iterations (lambda
    1e-5) with FP32 products and FP64 vectors. Reports kernel TFLOP/s (6d per
    pair) over every product launch of the solve, the dominant kernel's
    roofline, and the FP32 error against FP64 GPU products on a 100,000-point
    subsample whose FP64 results are checked against the oracle on a few rows
    and grid points."""
    import numpy as np

    import oracle
    from sgpp_amd import capi, workloads
    t_start = time.perf_counter()
    copy_reference(args.copy_reference_mib)

    def note(msg):
        print("[large] %s (%.0f s)" % (msg, time.perf_counter() - t_start), file=sys.stderr,
              flush=True)

    t = time.perf_counter()
    c = workloads.config_large_streaming(
        target=args.grid_target_large, M=args.data_points_large,
        progress=lambda k, n: note("refinement step %d: %d grid points" % (k, n)))
    gen_s = time.perf_counter() - t
    note("workload ready")
    lv, ix = c["grid"].level_index()
    X, y, lam = c["X"], c["y"], c["lam"]
    M, N, d = X.shape[0], lv.shape[0], lv.shape[1]
    t = time.perf_counter()
    op = capi.MultiEval(capi.BASIS_LINEAR, X, precision=capi.PREC_FP32)
    op.prepare(lv, ix)
    setup_s = time.perf_counter() - t
    note("operator prepared; 5-iteration CG")
    op.kernel_timing_enable(True)
    alpha, info = op.solve(y, lam, mode=capi.SOLVE_FP32, max_iterations=5, epsilon=0.0)
    nm, msm, nt, mst = op.kernel_timing()
    op.kernel_timing_enable(False)
    launches = nm + nt
    flop_launch = 6.0 * d * M * N
    kern_s = (msm + mst) * 1e-3
    dominant = "multT_kernel" if mst / nt >= msm / nm else "mult_kernel"
    dom_ms = max(mst / nt, msm / nm)
    # executed work of one launch of each kernel (counting instantiations, untimed)
    work = op.count_work(capi.PREC_FP32)
    key = "multT" if dominant == "multT_kernel" else "mult"
    executed = 4.0 * work["%s_hat_steps" % key] + 2.0 * work["%s_pairs" % key]
    note("CG done; subsample check")
    # subsample check: FP32 vs FP64 GPU products, FP64 vs the oracle
    S = min(M, 100000)
    sub32 = capi.MultiEval(capi.BASIS_LINEAR, X[:S], precision=capi.PREC_FP32).prepare(lv, ix)
    sub64 = capi.MultiEval(capi.BASIS_LINEAR, X[:S], precision=capi.PREC_FP64).prepare(lv, ix)
    a_test = workloads.normal(workloads.stream(6, 9), N)
    y_test = y[:S]
    r32, r64 = sub32.mult(a_test), sub64.mult(a_test)
    t32, t64 = sub32.mult_transpose(y_test), sub64.mult_transpose(y_test)
    rows = np.arange(0, S, S // 64)[:64]
    cols = np.arange(0, N, N // 16)[:16]
    o_r = oracle.mult(0, X[rows], lv, ix, a_test)
    o_t = oracle.mult_transpose(0, X[:S], lv[cols], ix[cols], y_test)
    line = {
        "metric": "kernel TFLOP/s (6d per pair) over a 5-iteration CG, FP32 products",
        "value": launches * flop_launch / kern_s / 1e12,
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "fp32 products, fp64 accumulation and CG vectors",
        "data": "synthetic, seeded (splitmix64): adaptive Linear d=20 grid, Gaussian-mixture points",
        "config": {"workload": c["name"], "grid_points": N, "data_points": M, "dim": d,
                   "lambda": lam, "cg_iterations": info["iterations"],
                   "refinement_steps": c["steps"]},
        "pair_evals_per_s": launches * M * N / kern_s,
        "product_launches": {"mult": nm, "multT": nt},
        "kernel_ms_avg": {"mult": msm / nm, "multT": mst / nt},
        "solve_wall_s": info["seconds"],
        "roofline": {"bound": "valu", "kernel": dominant,
                     "achieved": executed / (dom_ms * 1e-3) / 1e12,
                     "peak": PEAK_TFLOPS["fp32"], "unit": "TFLOP/s",
                     "frac": executed / (dom_ms * 1e-3) / 1e12 / PEAK_TFLOPS["fp32"],
                     "executed_flops_per_launch": executed,
                     "hat_steps_per_launch": work["%s_hat_steps" % key],
                     "evaluated_pairs_per_launch": work["%s_pairs" % key],
                     "achieved_6d_convention": flop_launch / (dom_ms * 1e-3) / 1e12,
                     "flops_per_launch_6d_convention": flop_launch,
                     "traffic": None,
                     "flop_convention": "achieved/frac: flops the kernel executes (4 per hat "
                                        "step of an evaluated pair and dimension, 2 per "
                                        "evaluated pair), counted by the counting "
                                        "instantiation of the same kernel; "
                                        "achieved_6d_convention: 6*d flops for every one of "
                                        "the M*N pairs"},
        "executed_work": work,
        "subsample": {
            "points": S,
            "fp32_vs_fp64_mult_rel_l2": float(np.linalg.norm(r32 - r64) / np.linalg.norm(r64)),
            "fp32_vs_fp64_multT_rel_l2": float(np.linalg.norm(t32 - t64) / np.linalg.norm(t64)),
            "fp64_vs_oracle_mult_rel_l2": float(np.linalg.norm(r64[rows] - o_r) /
                                                np.linalg.norm(o_r)),
            "fp64_vs_oracle_multT_rel_l2": float(np.linalg.norm(t64[cols] - o_t) /
                                                 np.linalg.norm(o_t)),
            "oracle_sample": "%d rows (mult), %d grid points (multTranspose)" % (len(rows),
                                                                                 len(cols)),
        },
        "generation_s": gen_s,
        "setup_s": setup_s,
    }
    print(json.dumps(line))


def copy_reference(mib):
    """mib MiB copied device to device three times: a kernel of known traffic
    (__amd_rocclr_copyBuffer) for scaling the hardware byte counters"""
    if mib <= 0:
        return
    import torch
    src = torch.ones(mib << 18, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU only: every box of this build has one GPU, "
              "so --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.serve:
        serve_bench(args)
        return
    if args.solve:
        solve_bench(args)
        return
    if args.large:
        large_bench(args)
        return
    import numpy as np
    import torch

    from sgpp_amd import capi, workloads

    t_setup = time.perf_counter()
    c = workloads.config_modlinear_adaptive(target=args.grid_target, M=args.data_points)
    lv, ix = c["grid"].level_index()
    X, alpha, y = c["X"], c["alpha"], c["y"]
    M, N, d = X.shape[0], lv.shape[0], lv.shape[1]
    dev = torch.device("cuda:0")
    d_alpha = torch.tensor(alpha, device=dev)
    d_y = torch.tensor(y, device=dev)
    d_r = torch.empty(M, dtype=torch.float64, device=dev)
    d_a = torch.empty(N, dtype=torch.float64, device=dev)
    op = capi.MultiEval(capi.BASIS_MODLINEAR, X, precision=capi.PREC_FP32, keep_fp64=True)
    op.prepare(lv, ix)
    setup_s = time.perf_counter() - t_setup
    copy_reference(args.copy_reference_mib)

    res = {}
    outs = {}
    for name, prec in (("fp64", capi.PREC_FP64), ("fp32", capi.PREC_FP32)):
        step_s, mult_ms, multT_ms = timed_steps(op, prec, d_alpha, d_y, d_r, d_a, args.steps,
                                                args.warmup, torch)
        res[name] = dict(step_s=step_s, mult_ms=mult_ms, multT_ms=multT_ms,
                         shape=op.last_shape())
        outs[name] = (d_r.cpu().numpy().copy(), d_a.cpu().numpy().copy())
    # executed work of each kernel (counting instantiations, untimed)
    work = {name: op.count_work(prec) for name, prec in (("fp64", capi.PREC_FP64),
                                                          ("fp32", capi.PREC_FP32))}

    pairs = float(M) * float(N)
    flops_6d = 6.0 * d * pairs
    v32 = 2.0 * pairs / res["fp32"]["step_s"]
    v64 = 2.0 * pairs / res["fp64"]["step_s"]

    def roof(name, peak):
        k = res[name]
        dom = "mult_kernel" if k["mult_ms"] >= k["multT_ms"] else "multT_kernel"
        key = "mult" if dom == "mult_kernel" else "multT"
        dom_ms = k["%s_ms" % key]
        w = work[name]
        executed = 4.0 * w["%s_hat_steps" % key] + 2.0 * w["%s_pairs" % key]
        return dom, dom_ms, {
            "achieved": executed / (dom_ms * 1e-3) / 1e12, "peak": peak,
            "frac": executed / (dom_ms * 1e-3) / 1e12 / peak,
            "executed_flops_per_launch": executed,
            "hat_steps_per_launch": w["%s_hat_steps" % key],
            "evaluated_pairs_per_launch": w["%s_pairs" % key],
            "achieved_6d_convention": flops_6d / (dom_ms * 1e-3) / 1e12,
        }

    dom, dom_ms, r32 = roof("fp32", PEAK_TFLOPS["fp32"])
    dom64, dom64_ms, r64 = roof("fp64", PEAK_TFLOPS["fp64"])
    traffic, tinfo = profile_traffic(dom, "float")
    k32 = res["fp32"]
    line = {
        "metric": "pair-evals/s (mult+multTranspose) vs FP64",
        "value": v32,
        "unit": "pair-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["fp32"]["step_s"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 products, fp64 accumulation",
        "data": "synthetic, seeded (splitmix64): adaptive ModLinear grid from the refinement loop, "
                "Gaussian-mixture data points",
        "config": {"workload": c["name"], "basis": "modlinear", "dim": d, "grid_points": N,
                   "data_points": M, "refinement_steps": c["steps"],
                   "launch_shape_fp32": {"mult_grid_chunks": k32["shape"][0],
                                         "multT_data_segs": k32["shape"][1]}},
        "gflops_6d": 6.0 * d * v32 / 1e9,
        "fp64": {"value": v64, "unit": "pair-evals/s", "ms_per_step": res["fp64"]["step_s"] * 1e3,
                 "mult_ms": res["fp64"]["mult_ms"], "multT_ms": res["fp64"]["multT_ms"],
                 "gflops_6d": 6.0 * d * v64 / 1e9,
                 "roofline": dict(r64, bound="valu", kernel=dom64, unit="TFLOP/s")},
        "speedup_vs_fp64": v32 / v64,
        "kernel_ms_fp32": {"mult": k32["mult_ms"], "multT": k32["multT_ms"]},
        "executed_work": work,
        "roofline": dict(r32, bound="valu", kernel=dom, unit="TFLOP/s", traffic=traffic,
                         traffic_source=tinfo,
                         flop_convention="achieved/frac: flops the kernel executes, counted in "
                                         "this run by the counting instantiation of the same "
                                         "kernel (4 per hat step of an evaluated pair and "
                                         "dimension, 2 per evaluated pair), over the kernel's "
                                         "HIP-event time; achieved_6d_convention: 6*d flops for "
                                         "every one of the M*N pairs "
                                         "(LearnerVectorizedPerformanceCalculator), most of which "
                                         "support culling skips as exact zeros"),
        "setup_s": setup_s,
    }
    # the subspace-indexed operator (FP64) on the same workload
    sop = capi.MultiEval(capi.BASIS_MODLINEAR, X, algorithm=capi.ALGO_SUBSPACE).prepare(lv, ix)
    sstep, smult, smultT = timed_steps(sop, capi.PREC_FP64, d_alpha, d_y, d_r, d_a, args.steps,
                                       args.warmup, torch)
    rs, as_ = d_r.cpu().numpy(), d_a.cpu().numpy()
    line["subspace_fp64"] = {
        "value": 2.0 * pairs / sstep, "unit": "pair-evals/s", "ms_per_step": sstep * 1e3,
        "mult_ms": smult, "multT_ms": smultT, "speedup_vs_streaming_fp64": res["fp64"]["step_s"] / sstep,
        "executed_work": sop.count_work(capi.PREC_FP64),
        "rel_diff_vs_streaming_fp64": {
            "mult": float(np.linalg.norm(rs - outs["fp64"][0]) / np.linalg.norm(outs["fp64"][0])),
            "multT": float(np.linalg.norm(as_ - outs["fp64"][1]) / np.linalg.norm(outs["fp64"][1]))},
    }
    del sop
    if not args.no_check:
        line["rel_err"] = accuracy(capi.BASIS_MODLINEAR, X, lv, ix, alpha, y, outs)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(capi.BASIS_MODLINEAR, X, c["grid"], lv, ix, alpha, y,
                                            args.cpu_seconds)
    del op, d_alpha, d_y, d_r, d_a
    if not args.no_serve:
        leg, _ = serve_leg([2000.0, 16000.0], args.serve_leg_seconds)
        line["served"] = leg
    print(json.dumps(line))


def profile_traffic(kernel, dtype):
    """HBM bytes per launch of the kernel from the committed rocprofv3 counter
    passes (profiles/traffic_<kernel>.json), used only when the profile was
    taken of the same kernel sources (SHA-256 over the device sources)."""
    import hashlib
    path = os.path.join(ROOT, "profiles", "traffic_%s.json" % kernel)
    h = hashlib.sha256()
    for f in KERNEL_SOURCES:
        with open(os.path.join(ROOT, "sgpp_amd", "csrc", f), "rb") as fh:
            h.update(fh.read())
    digest = h.hexdigest()
    if not os.path.exists(path):
        return None, {"file": None, "sources_sha256": digest, "match": False}
    with open(path) as f:
        tj = json.load(f)
    match = tj.get("sources_sha256") == digest
    return (tj.get("bytes_per_launch") if match else None), {
        "file": os.path.relpath(path, ROOT), "sources_sha256": digest,
        "profile_sources_sha256": tj.get("sources_sha256"), "match": match,
        "profile_kernel": tj.get("kernel")}


KERNEL_SOURCES = ("meval_kernels.hpp", "cull_device.hpp", "meval_kernels.hip")


if __name__ == "__main__":
    main()
