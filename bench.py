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
cpu_baseline: the reference's StreamingDefault ModLinear operator
(OperationMultiEvalModMaskStreaming, compiled from the reference sources into
oracle/_ref) on the host cores, FP64, on a bounded data sample.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu]
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


def cpu_baseline(basis, X, grid, lv, ix, alpha, y, seconds):
    import numpy as np

    import oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or oracle.num_threads()
    N = lv.shape[0]
    kind = "reference" if oracle.ref() is not None else "port"
    # the reference operator reads the same points, inserted in the same order
    rgrid = oracle.RefGrid.from_arrays(basis, lv, ix) if kind == "reference" else None

    def run(m):
        if kind == "reference":
            op = oracle.RefOp(rgrid, oracle.KIND_STREAMING, X[:m])
            op.mult(alpha)  # warm-up, first-touch
            t = time.perf_counter()
            op.mult(alpha)
            op.mult_transpose(y[:m])
            return time.perf_counter() - t
        op = oracle.Port(basis, X[:m], lv, ix)
        op.mult(alpha)
        t = time.perf_counter()
        op.mult(alpha)
        op.mult_transpose(y[:m])
        return time.perf_counter() - t

    m0 = 20000
    t0 = run(m0)
    m = int(min(X.shape[0], max(m0, m0 * seconds / max(t0, 1e-3))))
    times = sorted(run(m) for _ in range(3))
    med = times[1]
    return {
        "value": 2.0 * m * N / med,
        "unit": "pair-evals/s",
        "cores": threads,
        "kind": kind,
        "sample": ("%s StreamingDefault ModLinear operator (OperationMultiEvalModMaskStreaming%s), "
                   "FP64, full %d-point grid x first %d data points, mult+multTranspose, median of 3"
                   % ("reference" if kind == "reference" else "restated",
                      ", %s build" % oracle.ref_variant() if kind == "reference" else "", N, m)),
    }


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU only: every box of this build has one GPU, "
              "so --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)
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

    res = {}
    outs = {}
    for name, prec in (("fp64", capi.PREC_FP64), ("fp32", capi.PREC_FP32)):
        step_s, mult_ms, multT_ms = timed_steps(op, prec, d_alpha, d_y, d_r, d_a, args.steps,
                                                args.warmup, torch)
        res[name] = dict(step_s=step_s, mult_ms=mult_ms, multT_ms=multT_ms,
                         shape=op.last_shape())
        outs[name] = (d_r.cpu().numpy().copy(), d_a.cpu().numpy().copy())

    pairs = float(M) * float(N)
    flops_per_launch = 6.0 * d * pairs
    v32 = 2.0 * pairs / res["fp32"]["step_s"]
    v64 = 2.0 * pairs / res["fp64"]["step_s"]
    k32 = res["fp32"]
    dom_ms = max(k32["mult_ms"], k32["multT_ms"])
    dom = "mult_kernel" if k32["mult_ms"] >= k32["multT_ms"] else "multT_kernel"
    achieved = flops_per_launch / (dom_ms * 1e-3) / 1e12
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_%s.json" % dom)
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("bytes_per_launch")

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
        "gflops": 6.0 * d * v32 / 1e9,
        "fp64": {"value": v64, "unit": "pair-evals/s", "ms_per_step": res["fp64"]["step_s"] * 1e3,
                 "mult_ms": res["fp64"]["mult_ms"], "multT_ms": res["fp64"]["multT_ms"],
                 "gflops": 6.0 * d * v64 / 1e9,
                 "tflops_vs_fp64_peak": 6.0 * d * pairs / (max(res["fp64"]["mult_ms"], res["fp64"]["multT_ms"]) * 1e-3) / 1e12 / PEAK_TFLOPS["fp64"]},
        "speedup_vs_fp64": v32 / v64,
        "kernel_ms_fp32": {"mult": k32["mult_ms"], "multT": k32["multT_ms"]},
        "roofline": {"bound": "valu", "kernel": dom, "achieved": achieved,
                     "peak": PEAK_TFLOPS["fp32"], "unit": "TFLOP/s",
                     "frac": achieved / PEAK_TFLOPS["fp32"], "traffic": traffic,
                     "flops_per_launch": flops_per_launch,
                     "flop_convention": "6*d per pair (LearnerVectorizedPerformanceCalculator)"},
        "setup_s": setup_s,
    }
    if not args.no_check:
        line["rel_err"] = accuracy(capi.BASIS_MODLINEAR, X, lv, ix, alpha, y, outs)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(capi.BASIS_MODLINEAR, X, c["grid"], lv, ix, alpha, y,
                                            args.cpu_seconds)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
