"""The oracle itself is pinned here (CPU only): closed forms of the bases, the
reference's known-answer tests, the BT/BBT golden matrices, the reference's
DEFAULT outputs on the Friedman datasets and, where it was built, the
compiled reference operators. A test that plants a mistake shows the checks
would catch it."""
import json
import os

import numpy as np
import pytest

import oracle
from sgpp_amd import capi, workloads

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
ref_only = pytest.mark.skipif(oracle.ref() is None, reason="reference build not available")


def test_linear_basis_closed_form():
    # hat: 1 at the grid point, 0 at the support ends, linear in between
    for l in range(1, 8):
        for i in range(1, 2 ** l, 2):
            h = 2.0 ** -l
            assert oracle.phi(0, l, i, i * h) == 1.0
            assert oracle.phi(0, l, i, (i - 1) * h) == 0.0
            assert oracle.phi(0, l, i, (i + 1) * h) == 0.0
            assert abs(oracle.phi(0, l, i, (i - 0.5) * h) - 0.5) < 1e-15
            assert oracle.phi(0, l, i, min(1.0, (i + 2) * h)) == 0.0


def test_modlinear_basis_closed_form():
    assert oracle.phi(1, 1, 1, 0.0) == 1.0 and oracle.phi(1, 1, 1, 0.77) == 1.0
    for l in range(2, 8):
        h = 2.0 ** -l
        # left: 2 - 2^l x on [0, 2h], 0 beyond
        assert oracle.phi(1, l, 1, 0.0) == 2.0
        assert oracle.phi(1, l, 1, h) == 1.0
        assert oracle.phi(1, l, 1, 2 * h) == 0.0
        assert oracle.phi(1, l, 1, 3 * h) == 0.0
        # right: mirror image
        r = 2 ** l - 1
        assert oracle.phi(1, l, r, 1.0) == 2.0
        assert oracle.phi(1, l, r, 1.0 - h) == 1.0
        assert oracle.phi(1, l, r, 1.0 - 2 * h) == 0.0
        if l >= 3:  # interior functions are hats
            assert oracle.phi(1, l, 3, 3 * h) == 1.0
            assert oracle.phi(1, l, 3, 2 * h) == 0.0


KAT_POINTS = np.array([[0.5, 1.0], [0.3, 0.4], [0.9, 0.7]])


def kat_2d(level_index):
    lv, ix = level_index
    alpha = np.arange(1, lv.shape[0] + 1, dtype=np.float64)
    return oracle.mult(0, KAT_POINTS, lv, ix, alpha)


def test_kat_bounding_box_example():
    # base/tests/test_OperationMultipleEval.cpp:18-72 (points already mapped to [0,1]^2)
    r = kat_2d(capi.Grid.linear(2).regular(2).level_index())
    assert np.allclose(r, [0.0, 2.72, 1.64], rtol=1e-9, atol=1e-12)


def test_kat_planted_mistake_is_caught():
    # swapping the two dimensions of the grid (a transposed operand) breaks the KAT
    lv, ix = capi.Grid.linear(2).regular(2).level_index()
    r = kat_2d((lv[:, ::-1].copy(), ix[:, ::-1].copy()))
    assert not np.allclose(r, [0.0, 2.72, 1.64], rtol=1e-9, atol=1e-12)
    # a wrong index (i -> i + 2) breaks it too
    r = kat_2d((lv, ix + 2))
    assert not np.allclose(r, [0.0, 2.72, 1.64], rtol=1e-9, atol=1e-12)


def test_kat_1d():
    # base/tests/test_GridFactory.cpp:409-438
    lv, ix = capi.Grid.linear(1).regular(2).level_index()
    x = np.array([[0.25]])
    a = oracle.mult_transpose(0, x, lv, ix, np.array([1.0]))
    assert a.tolist() == [0.5, 1.0, 0.0]
    r = oracle.mult(0, x, lv, ix, np.array([1.0, 0.0, 0.0]))
    assert r.tolist() == [0.5]


def _bt_cases():
    with open(os.path.join(GOLDEN, "bt_bbt.json")) as f:
        return json.load(f)


def dense_B(basis, X, lv, ix):
    N = lv.shape[0]
    B = np.empty((X.shape[0], N))
    for j in range(N):
        e = np.zeros(N)
        e[j] = 1.0
        B[:, j] = oracle.mult(basis, X, lv, ix, e)
    return B


@pytest.mark.parametrize("case", _bt_cases(), ids=lambda c: "b%d_d%d_l%d" % (c["basis"], c["dim"], c["level"]))
def test_bt_bbt_golden(case):
    X = np.array(case["data"])
    g = capi.Grid(case["basis"], case["dim"]).regular(case["level"])
    assert g.size == case["nops"]
    lv, ix = g.level_index()
    B = dense_B(case["basis"], X, lv, ix)
    # testBT.cpp:49-135 (BOOST_CHECK_CLOSE 1e-5 %)
    np.testing.assert_allclose(B.sum(1), case["bt_row_sums"], rtol=1e-7, atol=1e-12)
    np.testing.assert_allclose(sorted(B.sum(0)), case["bt_col_sums_sorted"], rtol=1e-7, atol=1e-12)
    # testBBT.cpp (1e-2 %)
    BB = B.T @ B
    np.testing.assert_allclose(sorted(np.diag(BB)), case["bbt_diag_sorted"], rtol=1e-4, atol=1e-9)
    np.testing.assert_allclose(sorted(BB.sum(1)), case["bbt_row_sums_sorted"], rtol=1e-4, atol=1e-9)
    np.testing.assert_allclose(sorted(BB.sum(0)), case["bbt_col_sums_sorted"], rtol=1e-4, atol=1e-9)


def _friedman():
    return np.load(os.path.join(GOLDEN, "friedman.npz"))


@pytest.mark.parametrize("tag", ["f2", "f1"])
@pytest.mark.parametrize("basis", [0, 1])
def test_friedman_default_outputs(tag, basis):
    z = _friedman()
    X = z[tag + "_X"]
    key = "%s_b%d" % (tag, basis)
    lv, ix = z[key + "_level"].astype(np.uint32), z[key + "_index"].astype(np.uint32)
    r = oracle.mult(basis, X, lv, ix, np.arange(lv.shape[0], dtype=np.float64))
    a = oracle.mult_transpose(basis, X, lv, ix, np.arange(1, X.shape[0] + 1, dtype=np.float64))
    assert np.max(np.abs(r - z[key + "_mult"])) <= 1e-12 * np.max(np.abs(z[key + "_mult"]))
    assert np.max(np.abs(a - z[key + "_multT"])) <= 1e-12 * np.max(np.abs(z[key + "_multT"]))


def test_friedman_grid_is_ours():
    # the fixture grid (regular 5 + one seeded surplus refinement) is what our engine builds
    z = _friedman()
    for tag, d in (("f2", 4), ("f1", 10)):
        for basis in (0, 1):
            g = capi.Grid(basis, d).regular(5)
            g.refine_surplus(1.0 + 99.0 * workloads.uniform(workloads.stream(7, basis), g.size), 80)
            lv, ix = g.level_index()
            key = "%s_b%d" % (tag, basis)
            assert (lv == z[key + "_level"]).all() and (ix == z[key + "_index"]).all()


@pytest.mark.parametrize("basis", [0, 1])
def test_port_matches_definition(basis):
    g = capi.Grid(basis, 4).regular(4)
    lv, ix = g.level_index()
    X = workloads.uniform_points(11, 1000, 4)
    alpha = workloads.normal(5, g.size)
    y = workloads.normal(6, 1000)
    p = oracle.Port(basis, X, lv, ix)
    r0, a0 = oracle.mult(basis, X, lv, ix, alpha), oracle.mult_transpose(basis, X, lv, ix, y)
    assert np.linalg.norm(p.mult(alpha) - r0) <= 1e-13 * np.linalg.norm(r0)
    assert np.linalg.norm(p.mult_transpose(y) - a0) <= 1e-13 * np.linalg.norm(a0)


@ref_only
@pytest.mark.parametrize("basis", [0, 1])
def test_reference_operators_match_definition(basis):
    d = 5
    g = oracle.RefGrid(basis, d).regular(4)
    lv, ix = g.level_index()
    X = workloads.uniform_points(12, 777, d)
    alpha = workloads.normal(7, g.size)
    y = workloads.normal(8, 777)
    r0, a0 = oracle.mult(basis, X, lv, ix, alpha), oracle.mult_transpose(basis, X, lv, ix, y)
    kinds = [oracle.KIND_DEFAULT, oracle.KIND_STREAMING] + ([oracle.KIND_SUBSPACE] if basis == 0 else [])
    for kind in kinds:
        op = oracle.RefOp(g, kind, X)
        assert np.linalg.norm(op.mult(alpha) - r0) <= 1e-13 * np.linalg.norm(r0), kind
        assert np.linalg.norm(op.mult_transpose(y) - a0) <= 1e-13 * np.linalg.norm(a0), kind
