"""The drop-in, proven with SG++'s own classes: the reference's
LearnerLeastSquaresIdentity::train (LearnerBase.cpp:113-327), its
SystemMatrixLeastSquaresIdentity (:18-65), ConjugateGradients and
ModelFittingLeastSquares, compiled from the reference sources into
oracle/_ref/learner_check, run with the HIP operator. The program's
op_factory::createOperationMultipleEval hands the HIP sub-type to the SG++-side
adapter (integration/OperationMultiEvalHIP.hpp, compiled against the
reference's headers) and passes every other configuration to the reference's
own factory.

Each run is compared with the same reference run on its StreamingDefault
operator. Tolerances: grids identical point for point; FP64 products: alpha,
predictions and train MSE within 1e-8 relative (CG to 1e-12); FP32 products
inside the reference's FP64 CG (which then converges to the FP32 system):
predictions within 1e-4 relative on a regular grid."""
import numpy as np
import pytest

import oracle
from sgpp_amd import workloads

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(oracle.learner_check_path() is None,
                                 reason="reference learner build not available")]


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def dataset(seed, M=3000, d=3):
    X = workloads.uniform_points(seed, M, d)
    y = np.sin(2 * np.pi * X[:, 0]) * X[:, 1] + X[:, 2] ** 2
    y += 0.01 * workloads.normal(workloads.stream(seed, 5), M)
    return X, y


@pytest.mark.parametrize("basis", [0, 1])
def test_reference_learner_on_the_hip_operator(gpu, basis):
    X, y = dataset(51)
    Xt, _ = dataset(52, M=400)
    cfg = dict(refine=(1e-12, 1000), num_refinements=3, num_points=5)
    ref, rinfo = oracle.run_learner_check("learner", "streaming", basis, X, y, 3, 1e-4, Xt=Xt,
                                          **cfg)
    hip, hinfo = oracle.run_learner_check("learner", "hip64", basis, X, y, 3, 1e-4, Xt=Xt,
                                          **cfg)
    # the system matrix's operator and predict's came from the HIP branch
    assert hinfo["hip_ops"] >= 2 and rinfo["hip_ops"] == 0
    rlv, rix, ra, rpred = ref["train"]
    lv, ix, a, pred = hip["train"]
    assert np.array_equal(lv, rlv) and np.array_equal(ix, rix)
    assert rel(a, ra) <= 1e-8
    assert rel(pred, rpred) <= 1e-8
    assert hinfo["train_mse"] == pytest.approx(rinfo["train_mse"], rel=1e-8)


def test_reference_learner_on_the_hip_subspace_operator(gpu):
    """The factory's SUBSPACELINEAR branch: the reference's learner on the
    adapter's subspace-indexed kernels against the same learner on the
    reference's own OperationMultipleEvalSubspaceCombined (Linear grids)."""
    X, y = dataset(55)
    Xt, _ = dataset(56, M=400)
    cfg = dict(refine=(1e-12, 1000), num_refinements=3, num_points=5)
    ref, rinfo = oracle.run_learner_check("learner", "subspace", 0, X, y, 3, 1e-4, Xt=Xt, **cfg)
    hip, hinfo = oracle.run_learner_check("learner", "hipsubspace", 0, X, y, 3, 1e-4, Xt=Xt,
                                          **cfg)
    assert hinfo["hip_ops"] >= 2 and rinfo["hip_ops"] == 0
    rlv, rix, ra, rpred = ref["train"]
    lv, ix, a, pred = hip["train"]
    assert np.array_equal(lv, rlv) and np.array_equal(ix, rix)
    assert rel(a, ra) <= 1e-8
    assert rel(pred, rpred) <= 1e-8
    assert hinfo["train_mse"] == pytest.approx(rinfo["train_mse"], rel=1e-8)


def test_reference_learner_fp32_products(gpu):
    X, y = dataset(53, d=4)
    Xt, _ = dataset(54, M=400, d=4)
    cfg = dict(refine=(1e-6, 1000), num_refinements=0)
    ref, _ = oracle.run_learner_check("learner", "streaming", 0, X, y, 4, 1e-4, Xt=Xt, **cfg)
    hip, info = oracle.run_learner_check("learner", "hip32", 0, X, y, 4, 1e-4, Xt=Xt, **cfg)
    assert info["hip_ops"] >= 2
    assert np.array_equal(hip["train"][0], ref["train"][0])
    assert rel(hip["train"][3], ref["train"][3]) <= 1e-4


@pytest.mark.parametrize("basis", [0, 1])
def test_reference_fitter_on_the_hip_operator(gpu, basis):
    X, y = dataset(55)
    Xt, _ = dataset(56, M=300)
    cg = (1e-12, 1000)
    kw = dict(Xt=Xt, refine=cg, final=cg, num_refinements=1, num_points=6)
    ref, _ = oracle.run_learner_check("fitter", "streaming", basis, X, y, 3, 1e-4, **kw)
    hip, info = oracle.run_learner_check("fitter", "hip64", basis, X, y, 3, 1e-4, **kw)
    assert info["hip_ops"] >= 4 and info["refined"] == 1
    for stage in ("fit", "adapt"):
        rlv, rix, ra, rpred = ref[stage]
        lv, ix, a, pred = hip[stage]
        assert np.array_equal(lv, rlv) and np.array_equal(ix, rix), stage
        assert rel(a, ra) <= 1e-8, stage
        assert rel(pred, rpred) <= 1e-8, stage


@pytest.mark.parametrize("basis,regularization", [(0, 0), (0, 1), (1, 0), (1, 2)])
def test_reference_regression_learner_on_the_hip_operator(gpu, basis, regularization):
    """RegressionLearner::train (RegressionLearner.cpp:62-128) with its
    DMSystemMatrix (DMSystemMatrix.cpp:16-86; Identity, Laplace or Diagonal
    regularization operator, which stays the reference's) and ConjugateGradients,
    all operators from the base factory: DEFAULT against the HIP adapter."""
    X, y = dataset(57, M=2000)
    Xt, _ = dataset(58, M=300)
    kw = dict(Xt=Xt, refine=(1e-12, 1000), num_refinements=2, num_points=4,
              regularization=regularization)
    ref, rinfo = oracle.run_learner_check("regression", "default", basis, X, y, 3, 1e-4, **kw)
    hip, hinfo = oracle.run_learner_check("regression", "hipdefault", basis, X, y, 3, 1e-4,
                                          **kw)
    assert hinfo["base_hip_ops"] > 10 and hinfo["base_reference_ops"] == 0
    assert rinfo["base_hip_ops"] == 0
    rlv, rix, _, rpred = ref["train"]
    lv, ix, _, pred = hip["train"]
    assert np.array_equal(lv, rlv) and np.array_equal(ix, rix)
    assert rel(pred, rpred) <= 1e-8
    assert hinfo["train_mse"] == pytest.approx(rinfo["train_mse"], rel=1e-8)


@pytest.mark.parametrize("basis", [0, 1])
def test_reference_weight_matrix_on_the_hip_operator(gpu, basis):
    """DMWeightMatrix (DMWeightMatrix.cpp:24-60): (B^T W B + lambda I) alpha =
    B^T W y in the reference's ConjugateGradients, DEFAULT against the HIP
    adapter."""
    X, y = dataset(59, M=2500, d=4)
    Xt, _ = dataset(60, M=200, d=4)
    w = 0.5 + X[:, 0] * X[:, 1]
    kw = dict(Xt=Xt, final=(1e-12, 2000), weights=w)
    ref, rinfo = oracle.run_learner_check("weighted", "default", basis, X, y, 4, 1e-4, **kw)
    hip, hinfo = oracle.run_learner_check("weighted", "hipdefault", basis, X, y, 4, 1e-4, **kw)
    assert hinfo["base_hip_ops"] == 2 and hinfo["base_reference_ops"] == 0
    ra, rpred = ref["train"][2], ref["train"][3]
    a, pred = hip["train"][2], hip["train"][3]
    assert rel(a, ra) <= 1e-8
    assert rel(pred, rpred) <= 1e-8
