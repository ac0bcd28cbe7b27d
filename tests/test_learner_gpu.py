"""Regression learner on the GPU operator against the reference's own
LearnerLeastSquaresIdentity::train (LearnerBase.cpp:113-327, compiled from
its sources into oracle/_ref/learner_check, with its StreamingDefault
operator), and per refinement pass against the same steps on reference parts
(oracle.learner_train, itself checked against the real learner in
tests/test_learner.py).

Tolerances: grids identical point for point after every refinement (the
refinement picks the same top-k surpluses); alpha and train MSE within 1e-8
relative for the FP64 learner (CG to 1e-12, so solver error stays below the
comparison). CG iteration counts depend on rounding once the system is
ill-conditioned: the reference's own DEFAULT and StreamingDefault operators
differ by up to 5 of 290 iterations on the ModLinear case here, so counts are
asserted within 10 %; the mixed-precision learner
stops on ||r|| <= eps ||b||, so only its grid and a 1e-6 alpha agreement are
asserted."""
import numpy as np
import pytest

import oracle
from sgpp_amd import workloads

pytestmark = pytest.mark.gpu
ref_only = pytest.mark.skipif(oracle.ref() is None, reason="reference build not available")


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def dataset(seed=31, M=3000, d=3):
    X = workloads.uniform_points(seed, M, d)
    y = np.sin(2 * np.pi * X[:, 0]) * X[:, 1] + X[:, 2] ** 2
    y += 0.01 * workloads.normal(workloads.stream(seed, 5), M)
    return X, y


@ref_only
@pytest.mark.parametrize("basis", [0, 1])
def test_train_matches_reference_loop(gpu, basis):
    capi = gpu
    X, y = dataset()
    rg, ra, rsteps = oracle.learner_train(basis, X, y, 3, 1e-4, refine=(1e-12, 1000),
                                          num_refinements=4, num_points=5, test_acc=True)
    lr = capi.Learner(precision=capi.PREC_FP64)
    timing = lr.train(X, y, basis, 3, 1e-4, refine=(1e-12, 1000), num_refinements=4,
                      num_refinement_points=5, test_acc_during_adapt=True)
    steps = lr.steps()
    print([(s["grid_size"], s["iterations"]) for s in steps], rsteps, flush=True)
    assert len(steps) == len(rsteps)
    for s, (n, it, mse) in zip(steps, rsteps):
        assert s["grid_size"] == n
        assert abs(s["iterations"] - it) <= max(1, it // 10)
        assert s["train_accuracy"] == pytest.approx(mse, rel=1e-8)
    lv, ix = lr.level_index()
    rlv, rix = rg.level_index()
    assert np.array_equal(lv, rlv) and np.array_equal(ix, rix)
    assert rel(lr.alpha(), ra) <= 1e-8
    # predict == the definition of B alpha on new points
    Xt, yt = dataset(seed=32, M=500)
    p = lr.predict(Xt)
    assert rel(p, oracle.mult(basis, Xt, lv, ix, lr.alpha())) <= 1e-12
    assert lr.accuracy(Xt, yt) == pytest.approx(np.mean((p - yt) ** 2), rel=1e-12)
    # operation count of the reference's convention, summed over the passes;
    # refinement only appends, so pass k's grid is the first grid_size points
    gf = gb = 0.0
    for s in steps:
        n = s["grid_size"]
        f, b = oracle.RefGrid.from_arrays(basis, lv[:n], ix[:n]).performance(
            X.shape[0], s["iterations"], reuse_alpha=True)
        gf += f
        gb += b
    assert timing["gflop"] == pytest.approx(gf, rel=1e-12)
    assert timing["gbyte"] == pytest.approx(gb, rel=1e-12)
    assert timing["time_complete"] > 0
    assert timing["time_mult_compute"] > 0 and timing["time_mult_trans_compute"] > 0


learner_check = pytest.mark.skipif(oracle.learner_check_path() is None,
                                   reason="reference learner build not available")


@learner_check
@pytest.mark.parametrize("basis", [0, 1])
def test_train_matches_reference_learner(gpu, basis):
    """The final grid, alpha, predictions and train MSE of the GPU learner
    against the reference's LearnerLeastSquaresIdentity on the same data."""
    capi = gpu
    X, y = dataset(seed=34)
    Xt, yt = dataset(seed=35, M=400)
    cfg = dict(refine=(1e-12, 1000), num_refinements=3, num_points=5)
    ref, info = oracle.run_learner_check("learner", "streaming", basis, X, y, 3, 1e-4, Xt=Xt,
                                         **cfg)
    rlv, rix, ra, rpred = ref["train"]
    lr = capi.Learner(precision=capi.PREC_FP64)
    timing = lr.train(X, y, basis, 3, 1e-4, refine=cfg["refine"],
                      num_refinements=cfg["num_refinements"],
                      num_refinement_points=cfg["num_points"])
    lv, ix = lr.level_index()
    assert np.array_equal(lv, rlv) and np.array_equal(ix, rix)
    assert rel(lr.alpha(), ra) <= 1e-8
    assert rel(lr.predict(Xt), rpred) <= 1e-8
    assert lr.accuracy(X, y) == pytest.approx(info["train_mse"], rel=1e-8)
    # the reference's operation count depends on its CG iteration counts,
    # which rounding moves by a few iterations
    assert timing["gflop"] == pytest.approx(info["gflop"], rel=0.1)


@ref_only
def test_mixed_precision_learner(gpu):
    capi = gpu
    X, y = dataset(seed=33, M=4000, d=4)
    rg, ra, rsteps = oracle.learner_train(0, X, y, 3, 1e-4, refine=(1e-12, 1000),
                                          num_refinements=2, num_points=4)
    lr = capi.Learner(precision=capi.PREC_FP32)
    lr.train(X, y, 0, 3, 1e-4, refine=(1e-12, 1000), num_refinements=2, num_refinement_points=4)
    steps = lr.steps()
    assert [s["grid_size"] for s in steps] == [n for n, _, _ in rsteps]
    assert all(s["refinements"] >= 1 for s in steps)
    assert rel(lr.alpha(), ra) <= 1e-6


def test_learner_errors(gpu):
    capi = gpu
    X, y = dataset(M=200)
    lr = capi.Learner()
    with pytest.raises(capi.SGppHIPError) as e:
        lr.predict(X)
    assert e.value.status_name == "INVALID_ARGUMENT"
    with pytest.raises(ValueError):
        lr.train(X, y[:-1], 0, 2, 1e-3)
    with pytest.raises(capi.SGppHIPError) as e:
        lr.train(X, y, 0, 0, 1e-3)  # level 0
    assert e.value.status_name == "INVALID_ARGUMENT"
    with pytest.raises(capi.SGppHIPError) as e:
        lr.train(X, y, 0, 2, 1e-3, solver=capi.SLE_BICGSTAB)
    assert e.value.status_name == "INVALID_ARGUMENT"
    with pytest.raises(capi.SGppHIPError) as e:
        lr.train(X, y, 0, 2, 1e-3, error_based_refinement=True)
    assert e.value.status_name == "NOT_IMPLEMENTED"
    lr.train(X, y, 1, 2, 1e-3, num_refinements=0)
    assert lr.grid_size == 7
    with pytest.raises(capi.SGppHIPError):
        lr.predict(np.full((3, 3), 2.0))
