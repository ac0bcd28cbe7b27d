"""Learner operation counts (no GPU): the GFlop / GByte convention of the
reference's vectorised learners against the reference's own
LearnerVectorizedPerformanceCalculator compiled into oracle/_ref, plus one
count worked by hand."""
import numpy as np
import pytest

import oracle
from sgpp_amd import capi

ref_only = pytest.mark.skipif(oracle.ref() is None, reason="reference build not available")


def test_modlinear_1d_count_by_hand():
    # 1-D ModLinear level 2: (1,1) constant -> 0 flop, (2,1) left -> 8, (2,3) right -> 10;
    # 7 CG iterations, fresh start: 7 + 1.5 products; M = 1000 instances
    g = capi.Grid.modlinear(1).regular(2)
    gf, gb = g.performance(1000, 7, size_datatype=8, reuse_alpha=False)
    its = 8.5
    assert gf == pytest.approx(1e-9 * (8 + 10) * its * 1000, rel=1e-14)
    words = 1e-9 * (4 + 6) * its * 8 * 1000
    coeff = 1e-9 * its * (3 * 1001 * 8) + 1e-9 * its * (4 * 1000 * 8)
    assert gb == pytest.approx(words + coeff, rel=1e-14)


@ref_only
@pytest.mark.parametrize("basis", [0, 1])
@pytest.mark.parametrize("iterations,reuse,solver", [(0, False, 0), (37, True, 0),
                                                     (149, False, 0), (60, True, 1)])
def test_counts_match_reference(basis, iterations, reuse, solver):
    g = capi.Grid(basis, 4).regular(3)
    lv, ix = g.level_index()
    g.refine_surplus(np.linspace(1.0, 2.0, g.size), 6)
    lv, ix = g.level_index()
    rg = oracle.RefGrid.from_arrays(basis, lv, ix)
    mine = g.performance(12345, iterations, solver=solver, reuse_alpha=reuse)
    theirs = rg.performance(12345, iterations, solver=solver, reuse_alpha=reuse)
    assert mine[0] == pytest.approx(theirs[0], rel=1e-14)
    assert mine[1] == pytest.approx(theirs[1], rel=1e-14)


@pytest.mark.skipif(oracle.learner_check_path() is None or oracle.ref() is None,
                    reason="reference builds not available")
@pytest.mark.parametrize("basis", [0, 1])
def test_restated_learner_loop_matches_reference_learner(basis):
    """oracle.learner_train (LearnerBase::train's steps on reference parts,
    the per-pass oracle of the GPU learner tests) against the reference's own
    LearnerLeastSquaresIdentity::train: grids identical, alpha within 1e-8."""
    import numpy as np

    from sgpp_amd import workloads
    X = workloads.uniform_points(36, 2500, 3)
    y = np.sin(2 * np.pi * X[:, 0]) * X[:, 1] + X[:, 2] ** 2
    out, _ = oracle.run_learner_check("learner", "streaming", basis, X, y, 3, 1e-4,
                                      refine=(1e-12, 1000), num_refinements=3, num_points=5)
    rlv, rix, ra, _ = out["train"]
    g, a, _ = oracle.learner_train(basis, X, y, 3, 1e-4, refine=(1e-12, 1000), num_refinements=3,
                                   num_points=5)
    lv, ix = g.level_index()
    assert np.array_equal(lv, rlv) and np.array_equal(ix, rix)
    assert np.linalg.norm(a - ra) / np.linalg.norm(ra) <= 1e-8
