"""Regression solve (B^T B + lambda*M*I) alpha = B^T y on the GPU against the
reference's own ConjugateGradients + StreamingDefault (compiled into
oracle/_ref) and against the oracle's definition of B.

Tolerances: system matrix product 1e-12 relative (FP64); alpha of the FP64
GPU solve and of the mixed-precision (FP32 products, FP64 refinement) solve
within 1e-8 relative of the reference's FP64 CPU solve when all solves run
to a 1e-13 relative residual (1e-10 leaves ~kappa*1e-10 of solver error in
alpha, which the 1e-8 comparison is not meant to measure)."""
import numpy as np
import pytest

import oracle
from sgpp_amd import workloads

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def small_regression(capi, d=4, level=4, M=20000, seed=41):
    g = capi.Grid.linear(d).regular(level)
    X = workloads.uniform_points(seed, M, d)
    y = np.zeros(M)
    for j in range(d):
        y += np.sin(2 * np.pi * X[:, j]) * X[:, (j + 1) % d]
    y += 0.01 * workloads.normal(workloads.stream(seed, 5), M)
    return g, X, y


def test_system_matrix_product(gpu):
    capi = gpu
    g, X, y = small_regression(capi)
    lv, ix = g.level_index()
    lam = 1e-3
    v = workloads.normal(77, g.size)
    q0 = oracle.mult_transpose(0, X, lv, ix, oracle.mult(0, X, lv, ix, v)) + X.shape[0] * lam * v
    op = capi.MultiEval(0, X).prepare(lv, ix)
    assert rel(op.system_mult(v, lam), q0) <= 1e-12
    op32 = capi.MultiEval(0, X, precision=1, keep_fp64=True).prepare(lv, ix)
    assert rel(op32.system_mult(v, lam, precision=1), q0) <= 1e-5
    assert rel(op32.system_mult(v, lam, precision=0), q0) <= 1e-12


@pytest.mark.skipif(oracle.ref() is None, reason="reference build not available")
def test_cg_matches_reference_cg(gpu):
    capi = gpu
    g, X, y = small_regression(capi)
    lv, ix = g.level_index()
    lam, eps = 1e-4, 1e-13
    rg = oracle.RefGrid.from_arrays(0, lv, ix)
    rop = oracle.RefOp(rg, oracle.KIND_STREAMING, X)
    a_ref, it_ref, _ = rop.cg_solve(y, lam, 5000, eps)
    op = capi.MultiEval(0, X).prepare(lv, ix)
    a, info = op.solve(y, lam, mode=capi.SOLVE_FP64, max_iterations=5000, epsilon=eps)
    print("iterations gpu %d reference %d, residual %.2e, rel diff %.2e"
          % (info["iterations"], it_ref, info["residual_rel"], rel(a, a_ref)), flush=True)
    assert rel(a, a_ref) <= 1e-8
    assert abs(info["iterations"] - it_ref) <= max(3, it_ref // 20)
    # the same semantics at the reference's default tolerance: iteration counts agree
    a10, it10_ref, _ = rop.cg_solve(y, lam, 5000, 1e-10)
    b10, info10 = op.solve(y, lam, mode=capi.SOLVE_FP64, max_iterations=5000, epsilon=1e-10)
    assert abs(info10["iterations"] - it10_ref) <= max(2, it10_ref // 20)
    assert info10["residual_rel"] <= 2e-10


def test_mixed_precision_refinement(gpu):
    capi = gpu
    g, X, y = small_regression(capi, d=5, level=4, M=40000, seed=43)
    lv, ix = g.level_index()
    lam = 1e-4
    op = capi.MultiEval(0, X, precision=1, keep_fp64=True).prepare(lv, ix)
    a64, i64 = op.solve(y, lam, mode=capi.SOLVE_FP64, max_iterations=5000, epsilon=1e-13)
    amx, imx = op.solve(y, lam, mode=capi.SOLVE_MIXED, max_iterations=5000, epsilon=1e-13,
                        inner_epsilon=1e-5, max_refinements=30)
    print("fp64: it %d matvec64 %d  mixed: it %d refinements %d matvec64 %d matvec32 %d resid %.2e"
          % (i64["iterations"], i64["fp64_matvecs"], imx["iterations"], imx["refinements"],
             imx["fp64_matvecs"], imx["fp32_matvecs"], imx["residual_rel"]), flush=True)
    assert imx["residual_rel"] <= 1e-13 * 10
    assert imx["refinements"] >= 1 and imx["fp32_matvecs"] > 0
    assert imx["fp64_matvecs"] <= imx["refinements"] + 2
    assert rel(amx, a64) <= 1e-8
    # an FP64 operator cannot run the mixed solve
    op64 = capi.MultiEval(0, X).prepare(lv, ix)
    with pytest.raises(capi.SGppHIPError):
        op64.solve(y, lam, mode=capi.SOLVE_MIXED)


def test_regression_workload_against_reference_solve(gpu):
    """Linear d=8 level 6, 500,000 points, lambda 1e-4, epsilon 1e-10: the
    FP64 GPU solve and the mixed solve against alpha of the reference's own
    CPU solve (tests/golden/regression_alpha.npz, made by
    tests/golden/make_regression_fixture.py). The FP64 GPU solve runs the
    reference's CG recurrence, so its iteration count matches within 1; both
    alphas agree with the reference's to 1e-8 relative."""
    import os
    capi = gpu
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "regression_alpha.npz"))
    c = workloads.config_regression()
    lv, ix = c["grid"].level_index()
    op = capi.MultiEval(0, c["X"], precision=1, keep_fp64=True).prepare(lv, ix)
    a64, i64 = op.solve(c["y"], c["lam"], mode=capi.SOLVE_FP64, max_iterations=10000,
                        epsilon=1e-10)
    amx, imx = op.solve(c["y"], c["lam"], mode=capi.SOLVE_MIXED, max_iterations=10000,
                        epsilon=1e-10, inner_epsilon=1e-5, max_refinements=40)
    print("reference: it %d; fp64: it %d rel %.2e; mixed: refinements %d matvec32 %d rel %.2e"
          % (int(z["iterations"]), i64["iterations"], rel(a64, z["alpha"]), imx["refinements"],
             imx["fp32_matvecs"], rel(amx, z["alpha"])), flush=True)
    assert abs(i64["iterations"] - int(z["iterations"])) <= 1
    assert rel(a64, z["alpha"]) <= 1e-8
    assert imx["residual_rel"] <= 2e-10
    assert rel(amx, z["alpha"]) <= 1e-8


def test_fp32_products_cg(gpu):
    """The reference's CG on FP32 products (FP64 vectors): same iteration
    semantics, accuracy limited by the FP32 products. Checked against the
    FP64 solve: 1e-4 relative at a 1e-6 residual."""
    capi = gpu
    g, X, y = small_regression(capi, d=4, level=4, M=20000, seed=44)
    lv, ix = g.level_index()
    lam = 1e-4
    op32 = capi.MultiEval(0, X, precision=1).prepare(lv, ix)
    a32, i32 = op32.solve(y, lam, mode=capi.SOLVE_FP32, max_iterations=5000, epsilon=1e-6)
    op64 = capi.MultiEval(0, X).prepare(lv, ix)
    a64, i64 = op64.solve(y, lam, mode=capi.SOLVE_FP64, max_iterations=5000, epsilon=1e-13)
    print("fp32 CG: it %d resid %.2e rel %.2e" % (i32["iterations"], i32["residual_rel"],
                                                  rel(a32, a64)), flush=True)
    # start residual + one per iteration + the residual recomputed every 50 iterations
    it = i32["iterations"]
    assert i32["fp64_matvecs"] == 0 and i32["fp32_matvecs"] == 1 + it + (it - 1) // 50
    assert i32["residual_rel"] <= 1e-5
    assert rel(a32, a64) <= 1e-4
    # a fixed number of iterations (the large streaming case runs 5)
    a5, i5 = op32.solve(y, lam, mode=capi.SOLVE_FP32, max_iterations=5, epsilon=0.0)
    assert i5["iterations"] == 5 and i5["fp32_matvecs"] == 6
    # an FP32-only operator cannot run the FP64 solve
    with pytest.raises(capi.SGppHIPError):
        op32.solve(y, lam, mode=capi.SOLVE_FP64)


def test_mixed_mode_keeps_the_callers_iteration_budget(gpu):
    """max_iterations bounds all inner CG iterations of the mixed solve
    together (the budget a caller such as LearnerBase::train gives the
    reference's CG), and a solve that spends its budget before reaching the
    tolerance reports that it did not converge; with a full budget it does."""
    capi = gpu
    g, X, y = small_regression(capi)
    lv, ix = g.level_index()
    op = capi.MultiEval(0, X, precision=1, keep_fp64=True).prepare(lv, ix)
    for budget in (1, 5, 12):
        a, info = op.solve(y, 1e-4, mode=capi.SOLVE_MIXED, max_iterations=budget, epsilon=1e-12,
                           inner_epsilon=1e-3)
        assert info["iterations"] <= budget, (budget, info)
        assert not info["converged"] and info["residual_rel"] > 1e-12
    a, info = op.solve(y, 1e-4, mode=capi.SOLVE_MIXED, max_iterations=2000, epsilon=1e-12)
    assert info["converged"] and info["residual_rel"] <= 1e-12 and info["iterations"] <= 2000
    a64, info64 = op.solve(y, 1e-4, mode=capi.SOLVE_FP64, max_iterations=3, epsilon=1e-12)
    assert info64["iterations"] == 3 and not info64["converged"]
