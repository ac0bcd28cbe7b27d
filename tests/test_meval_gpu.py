"""Parity of the HIP operator against the oracle, through the C-ABI.

Tolerances (relative, ||got - ref||_2 / ||ref||_2 unless stated):
  FP64 kernel:                      1e-12
  FP32 with FP64 accumulation:      1e-5   (inputs representable in FP32)
"""
import json
import os

import numpy as np
import pytest

import oracle
from sgpp_amd import workloads

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TOL = {0: 1e-12, 1: 1e-5}


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def make_case(capi, basis, d, level, M, seed, fp32_inputs=False):
    g = capi.Grid(basis, d).regular(level)
    X = workloads.uniform_points(seed, M, d)
    if fp32_inputs:
        X = X.astype(np.float32).astype(np.float64)
    alpha = workloads.normal(workloads.stream(seed, 1), g.size)
    y = workloads.normal(workloads.stream(seed, 2), M)
    return g, X, alpha, y


@pytest.mark.parametrize("basis", [0, 1])
@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("d,level,M", [(1, 6, 300), (2, 5, 5000), (5, 5, 10000), (8, 3, 4097),
                                       (10, 3, 2000), (13, 2, 3000), (20, 2, 1500), (23, 2, 700)])
def test_parity_regular(gpu, basis, prec, d, level, M):
    capi = gpu
    g, X, alpha, y = make_case(capi, basis, d, level, M, seed=100 + d, fp32_inputs=(prec == 1))
    lv, ix = g.level_index()
    print("case basis=%d prec=%d d=%d level=%d M=%d N=%d" % (basis, prec, d, level, M, g.size),
          flush=True)
    op = capi.MultiEval(basis, X, precision=prec).prepare(lv, ix)
    r = op.mult(alpha)
    a = op.mult_transpose(y)
    r0 = oracle.mult(basis, X, lv, ix, alpha)
    a0 = oracle.mult_transpose(basis, X, lv, ix, y)
    assert rel(r, r0) <= TOL[prec], rel(r, r0)
    assert rel(a, a0) <= TOL[prec], rel(a, a0)


@pytest.mark.parametrize("basis", [0, 1])
@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("d,M", [(2, 3001), (3, 2500), (7, 2000), (12, 1500), (20, 800)])
def test_parity_adaptive(gpu, basis, prec, d, M):
    """Adaptive grids (regular level 2 or 3, then surplus refinement driven by
    a random indicator) whose points have varied sets of non-constant
    dimensions: exercises the ModLinear storage order, the per-tile and
    per-4-point masks and partly filled tiles."""
    capi = gpu
    g = capi.Grid(basis, d).regular(3 if d <= 7 else 2)
    for step in range(4 if d <= 7 else 3):
        g.refine_surplus(1.0 + workloads.uniform(workloads.stream(200 + d, step), g.size),
                         max(1, g.size // 6))
    lv, ix = g.level_index()
    X = workloads.uniform_points(300 + d, M, d)
    if prec == 1:
        X = X.astype(np.float32).astype(np.float64)
    alpha = workloads.normal(workloads.stream(300 + d, 1), g.size)
    y = workloads.normal(workloads.stream(300 + d, 2), M)
    print("case basis=%d prec=%d d=%d M=%d N=%d" % (basis, prec, d, M, g.size), flush=True)
    op = capi.MultiEval(basis, X, precision=prec).prepare(lv, ix)
    r = op.mult(alpha)
    a = op.mult_transpose(y)
    assert rel(r, oracle.mult(basis, X, lv, ix, alpha)) <= TOL[prec]
    assert rel(a, oracle.mult_transpose(basis, X, lv, ix, y)) <= TOL[prec]


def test_small_linear_workload(gpu):
    """Linear d=5 level 5 (N=1471), 10,000 uniform points: FP64 <= 1e-12 against
    the definition and against the reference's StreamingDefault / SubspaceLinear."""
    capi = gpu
    c = workloads.config_small_linear()
    lv, ix = c["grid"].level_index()
    r0 = oracle.mult(0, c["X"], lv, ix, c["alpha"])
    a0 = oracle.mult_transpose(0, c["X"], lv, ix, c["y"])
    op = capi.MultiEval(0, c["X"], precision=0).prepare(lv, ix)
    r, a = op.mult(c["alpha"]), op.mult_transpose(c["y"])
    assert rel(r, r0) <= 1e-12 and rel(a, a0) <= 1e-12
    assert np.max(np.abs(r - r0)) <= 1e-12 * np.max(np.abs(r0))
    if oracle.ref() is not None:
        rg = oracle.RefGrid(0, 5).regular(5)
        for kind in (oracle.KIND_STREAMING, oracle.KIND_SUBSPACE):
            rop = oracle.RefOp(rg, kind, c["X"])
            assert rel(r, rop.mult(c["alpha"])) <= 1e-12
            assert rel(a, rop.mult_transpose(c["y"])) <= 1e-12
    op32 = capi.MultiEval(0, c["X"], precision=1).prepare(lv, ix)
    assert rel(op32.mult(c["alpha"]), r0) <= 1e-5
    assert rel(op32.mult_transpose(c["y"]), a0) <= 1e-5


def test_operator_through_factory_and_refinement(gpu):
    """Operator bound to a grid through createOperationMultipleEval; prepare()
    re-reads the grid after refinement (OperationMultiEvalStreaming lifecycle)."""
    capi = gpu
    g = capi.Grid.modlinear(4).regular(3)
    X = workloads.uniform_points(21, 3000, 4)
    op = capi.MultiEval(1, X, grid=g).prepare()
    for step in range(3):
        lv, ix = g.level_index()
        alpha = workloads.normal(workloads.stream(22, step), g.size)
        y = workloads.normal(workloads.stream(23, step), X.shape[0])
        assert rel(op.mult(alpha), oracle.mult(1, X, lv, ix, alpha)) <= 1e-12
        assert rel(op.mult_transpose(y), oracle.mult_transpose(1, X, lv, ix, y)) <= 1e-12
        assert op.duration() > 0
        g.refine_surplus(np.abs(alpha), 10)
        op.prepare()


def test_known_answers(gpu):
    capi = gpu
    # base/tests/test_OperationMultipleEval.cpp:18-72
    g = capi.Grid.linear(2).regular(2)
    lv, ix = g.level_index()
    op = capi.MultiEval(0, np.array([[0.5, 1.0], [0.3, 0.4], [0.9, 0.7]])).prepare(lv, ix)
    np.testing.assert_allclose(op.mult(np.arange(1.0, 6.0)), [0.0, 2.72, 1.64], rtol=1e-12,
                               atol=1e-14)
    # base/tests/test_GridFactory.cpp:409-438
    g = capi.Grid.linear(1).regular(2)
    lv, ix = g.level_index()
    op = capi.MultiEval(0, np.array([[0.25]])).prepare(lv, ix)
    assert op.mult_transpose(np.array([1.0])).tolist() == [0.5, 1.0, 0.0]
    assert op.mult(np.array([1.0, 0.0, 0.0])).tolist() == [0.5]


def test_bt_bbt_golden(gpu):
    capi = gpu
    with open(os.path.join(GOLDEN, "bt_bbt.json")) as f:
        cases = json.load(f)
    for c in cases:
        X = np.array(c["data"])
        g = capi.Grid(c["basis"], c["dim"]).regular(c["level"])
        lv, ix = g.level_index()
        op = capi.MultiEval(c["basis"], X).prepare(lv, ix)
        N = g.size
        B = np.stack([op.mult(np.eye(N)[j]) for j in range(N)], axis=1)
        np.testing.assert_allclose(B.sum(1), c["bt_row_sums"], rtol=1e-7, atol=1e-12)
        np.testing.assert_allclose(sorted(B.sum(0)), c["bt_col_sums_sorted"], rtol=1e-7, atol=1e-12)
        BB = np.stack([op.mult_transpose(B[:, j]) for j in range(N)], axis=1)
        np.testing.assert_allclose(sorted(np.diag(BB)), c["bbt_diag_sorted"], rtol=1e-4, atol=1e-9)
        np.testing.assert_allclose(sorted(BB.sum(1)), c["bbt_row_sums_sorted"], rtol=1e-4, atol=1e-9)


@pytest.mark.parametrize("tag", ["f2", "f1"])
@pytest.mark.parametrize("basis", [0, 1])
def test_friedman_against_reference_default(gpu, tag, basis):
    """compareToReference / compareToReferenceTranspose (test_datadrivenCommon.cpp:205-341),
    with the reference's MSE thresholds (streaming*Test.cpp) and a relative bound."""
    capi = gpu
    z = np.load(os.path.join(GOLDEN, "friedman.npz"))
    X = z[tag + "_X"]
    key = "%s_b%d" % (tag, basis)
    lv, ix = z[key + "_level"].astype(np.uint32), z[key + "_index"].astype(np.uint32)
    op = capi.MultiEval(basis, X).prepare(lv, ix)
    r = op.mult(np.arange(lv.shape[0], dtype=np.float64))
    a = op.mult_transpose(np.arange(1, X.shape[0] + 1, dtype=np.float64))
    mse_r = np.mean((r - z[key + "_mult"]) ** 2)
    mse_a = np.mean((a - z[key + "_multT"]) ** 2)
    thr = {("f2", 0): (1e-24, 1e-19), ("f1", 0): (1e-21, 1e-26),
           ("f2", 1): (1e-22, 1e-17), ("f1", 1): (1e-16, 1e-20)}[(tag, basis)]
    print("mse mult %.3e (reference threshold %.0e), mse multT %.3e (%.0e)"
          % (mse_r, thr[0], mse_a, thr[1]), flush=True)
    assert rel(r, z[key + "_mult"]) <= 1e-12 and rel(a, z[key + "_multT"]) <= 1e-12
    assert mse_r < thr[0] and mse_a < thr[1]


def test_edge_cases(gpu):
    capi = gpu
    # one data point, one grid point
    g = capi.Grid.linear(3).regular(1)
    lv, ix = g.level_index()
    op = capi.MultiEval(0, np.array([[0.5, 0.25, 0.75]])).prepare(lv, ix)
    assert op.mult(np.array([2.0])).tolist() == [2.0 * 1.0 * 0.5 * 0.5]
    assert op.mult_transpose(np.array([3.0])).tolist() == [3.0 * 0.25]
    # points on the boundary of the unit cube and of supports
    X = np.array([[0.0, 0.0], [1.0, 1.0], [0.5, 0.5], [0.25, 0.75], [0.125, 1.0]])
    for basis in (0, 1):
        g = capi.Grid(basis, 2).regular(4)
        lv, ix = g.level_index()
        alpha = np.linspace(-1, 1, g.size)
        op = capi.MultiEval(basis, X).prepare(lv, ix)
        np.testing.assert_allclose(op.mult(alpha), oracle.mult(basis, X, lv, ix, alpha),
                                   rtol=1e-13, atol=1e-14)
    # invalid inputs are rejected with a status, not computed
    with pytest.raises(capi.SGppHIPError) as e:
        capi.MultiEval(0, np.array([[1.5, 0.2]]))
    assert e.value.status_name == "INVALID_ARGUMENT"
    op = capi.MultiEval(0, X)
    with pytest.raises(capi.SGppHIPError):
        op.mult_transpose(np.ones(X.shape[0]))  # no grid prepared
    with pytest.raises(capi.SGppHIPError):
        op.prepare(np.zeros((3, 2)), np.ones((3, 2)))  # level 0 is not a no-boundary point


@pytest.mark.parametrize("basis", [0, 1])
@pytest.mark.parametrize("prec", [0, 1])
def test_culling_is_exact(gpu, basis, prec):
    """Support culling drops only exact zeros: data on support boundaries
    (dyadic coordinates, including 0 and 1), a box of zero width (one point
    repeated) and a tight cluster beside uniform points, each checked against
    the oracle, and a point outside every support reading exactly 0."""
    capi = gpu
    tol = 1e-13 if prec == 0 else 1e-6
    rng = np.random.default_rng(7)
    d = 3
    g = capi.Grid(basis, d).regular(5)
    lv, ix = g.level_index()
    alpha = workloads.normal(workloads.stream(41, 1), g.size)
    dyadic = rng.integers(0, 65, size=(5000, d)) / 64.0
    same = np.tile(np.array([[0.3125, 0.6875, 0.5]]), (3000, 1))
    cluster = np.vstack([0.8 + 0.001 * rng.random((2500, d)), rng.random((2500, d))])
    if prec == 1:
        cluster = cluster.astype(np.float32).astype(np.float64)
    for X in (dyadic, same, cluster):
        y = workloads.normal(workloads.stream(42, X.shape[0]), X.shape[0])
        op = capi.MultiEval(basis, X, precision=prec).prepare(lv, ix)
        r, a = op.mult(alpha), op.mult_transpose(y)
        r0 = oracle.mult(basis, X, lv, ix, alpha)
        a0 = oracle.mult_transpose(basis, X, lv, ix, y)
        assert rel(r, r0) <= tol and rel(a, a0) <= tol
        assert np.max(np.abs(r - r0)) <= tol * np.max(np.abs(r0))
    if basis == 0:
        # every Linear basis function vanishes at the corner x = (0, 0, 0), so
        # every multTranspose entry is an exact zero
        X = np.zeros((64, d))
        op = capi.MultiEval(basis, X, precision=prec).prepare(lv, ix)
        a = op.mult_transpose(np.ones(64))
        assert np.all(a == 0.0)


def test_launch_shapes_agree_and_are_deterministic(gpu):
    capi = gpu
    g, X, alpha, y = make_case(capi, 1, 6, 4, 20000, seed=31)
    lv, ix = g.level_index()
    base = capi.MultiEval(1, X).prepare(lv, ix)
    r0, a0 = base.mult(alpha), base.mult_transpose(y)
    assert (base.mult(alpha) == r0).all() and (base.mult_transpose(y) == a0).all()
    for G, S in [(1, 1), (2, 3), (7, 16)]:
        op = capi.MultiEval(1, X, mult_grid_chunks=G, multT_data_segs=S).prepare(lv, ix)
        assert rel(op.mult(alpha), r0) <= 1e-14 and rel(op.mult_transpose(y), a0) <= 1e-14


def test_device_pointers_and_adjoint_identity(gpu):
    import torch
    capi = gpu
    c = workloads.config_modlinear_adaptive(target=30000, M=200000)
    lv, ix = c["grid"].level_index()
    dev = torch.device("cuda:0")
    alpha = torch.tensor(c["alpha"], device=dev)
    y = torch.tensor(c["y"], device=dev)
    r = torch.empty(c["X"].shape[0], dtype=torch.float64, device=dev)
    a = torch.empty(lv.shape[0], dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    for prec, tol in ((0, 1e-12), (1, 1e-5)):
        op = capi.MultiEval(1, c["X"], precision=prec).prepare(lv, ix)
        op.mult_device(alpha.data_ptr(), r.data_ptr())
        op.mult_transpose_device(y.data_ptr(), a.data_ptr())
        op.synchronize()
        # <B alpha, y> == <alpha, B^T y>
        lhs = float(torch.dot(r, y))
        rhs = float(torch.dot(alpha, a))
        scale = float(torch.linalg.norm(r) * torch.linalg.norm(y))
        assert abs(lhs - rhs) <= tol * scale
        assert op.kernel_ms() > 0


def test_full_size_properties(gpu):
    """The full ModLinear d=10 workload (N = 100,015 adaptive grid points,
    M = 1,000,000 Gaussian-mixture points), checked through properties that
    hold at any size: the adjoint identity <B a, y> = <a, B^T y>, linearity
    B(a1 + 2 a2) = B a1 + 2 B a2 and the same for B^T, FP32 against FP64, and
    bit-identical repeats. Sampled rows and columns are also checked against
    the oracle."""
    import torch
    capi = gpu
    c = workloads.config_modlinear_adaptive()
    lv, ix = c["grid"].level_index()
    X = c["X"]
    M, N = X.shape[0], lv.shape[0]
    assert (M, N) == (1000000, 100015)
    dev = torch.device("cuda:0")
    a1 = torch.tensor(c["alpha"], device=dev)
    a2 = torch.tensor(workloads.normal(workloads.stream(77, 1), N), device=dev)
    y1 = torch.tensor(c["y"], device=dev)
    y2 = torch.tensor(workloads.normal(workloads.stream(77, 2), M), device=dev)
    op = capi.MultiEval(1, X, precision=1, keep_fp64=True).prepare(lv, ix)

    def mult(a, prec):
        r = torch.empty(M, dtype=torch.float64, device=dev)
        op.mult_device(a.data_ptr(), r.data_ptr(), prec)
        op.synchronize()
        return r

    def mult_t(y, prec):
        r = torch.empty(N, dtype=torch.float64, device=dev)
        op.mult_transpose_device(y.data_ptr(), r.data_ptr(), prec)
        op.synchronize()
        return r

    def rel_t(a, b):
        return float(torch.linalg.norm(a - b) / torch.linalg.norm(b))

    r64, t64 = mult(a1, 0), mult_t(y1, 0)
    # adjoint identity
    lhs, rhs = float(torch.dot(r64, y1)), float(torch.dot(a1, t64))
    assert abs(lhs - rhs) <= 1e-12 * float(torch.linalg.norm(r64) * torch.linalg.norm(y1))
    # linearity
    assert rel_t(mult(a1 + 2 * a2, 0), r64 + 2 * mult(a2, 0)) <= 1e-12
    assert rel_t(mult_t(y1 + 2 * y2, 0), t64 + 2 * mult_t(y2, 0)) <= 1e-12
    # repeats are bit-identical
    assert torch.equal(mult(a1, 0), r64) and torch.equal(mult_t(y1, 0), t64)
    assert torch.equal(mult(a1, 1), mult(a1, 1)) and torch.equal(mult_t(y1, 1), mult_t(y1, 1))
    # FP32 products with FP64 accumulation against FP64
    assert rel_t(mult(a1, 1), r64) <= 1e-5 and rel_t(mult_t(y1, 1), t64) <= 1e-5
    # sampled rows / columns against the oracle
    rows = np.arange(0, M, M // 97)[:97]
    cols = np.arange(0, N, N // 13)[:13]
    r0 = oracle.mult(1, X[rows], lv, ix, c["alpha"])
    t0 = oracle.mult_transpose(1, X, lv[cols], ix[cols], c["y"])
    assert rel(r64.cpu().numpy()[rows], r0) <= 1e-12
    assert rel(t64.cpu().numpy()[cols], t0) <= 1e-12


def test_range_forms(gpu):
    """mult / multTranspose over a range of data rows / grid points
    (OperationMultipleEval.hpp:70-73, 97-100): the slice of the full product,
    for an operator from the factory and for a plain handle, both precisions
    of the streaming kernels and the subspace kernels; empty and invalid
    ranges."""
    capi = gpu
    g = capi.Grid.modlinear(3).regular(4)
    lv, ix = g.level_index()
    X = workloads.uniform_points(61, 3000, 3)
    alpha = workloads.normal(workloads.stream(61, 1), g.size)
    y = workloads.normal(workloads.stream(61, 2), X.shape[0])
    ops = [capi.MultiEval(1, X, grid=g).prepare(),
           capi.MultiEval(1, X, precision=1).prepare(lv, ix),
           capi.MultiEval(1, X, algorithm=capi.ALGO_SUBSPACE).prepare(lv, ix)]
    for op in ops:
        r, a = op.mult(alpha), op.mult_transpose(y)
        assert np.array_equal(op.mult_range(alpha, 100, 2100), r[100:2100])
        assert np.array_equal(op.mult_range(alpha, 0, X.shape[0]), r)
        # (the subspace multTranspose sums with atomics: equal to rounding)
        assert rel(op.mult_transpose_range(y, 7, g.size), a[7:]) <= 1e-14
        assert op.mult_range(alpha, 5, 5).size == 0
        for bad in ((10, 5), (0, X.shape[0] + 1)):
            with pytest.raises(capi.SGppHIPError) as e:
                op.mult_range(alpha, *bad)
            assert e.value.status_name == "INVALID_ARGUMENT"
        with pytest.raises(capi.SGppHIPError) as e:
            op.mult_transpose_range(y, 0, g.size + 1)
        assert e.value.status_name == "INVALID_ARGUMENT"
