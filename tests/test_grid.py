"""Grid engine (input builder): sequence numbers, leaf flags and surpluses
must equal SG++'s bit for bit. CPU only."""
import math
import os

import numpy as np
import pytest

import oracle
from sgpp_amd import capi, workloads

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

ref_only = pytest.mark.skipif(oracle.ref() is None, reason="reference build not available")


def regular_size(d, L):
    # number of inner points of the regular sparse grid: sum_{k<L} 2^k C(d-1+k, d-1)
    return sum(2 ** k * math.comb(d - 1 + k, d - 1) for k in range(L))


@pytest.mark.parametrize("d,L", [(1, 1), (1, 5), (2, 3), (3, 4), (5, 5), (8, 6), (10, 4), (20, 2)])
def test_regular_size_closed_form(d, L):
    g = capi.Grid.linear(d).regular(L)
    assert g.size == regular_size(d, L)


def test_known_sizes():
    # test_GridFactory.cpp:379-391 and the sizes the BT_BBT fixtures use
    assert capi.Grid.linear(2).regular(3).size == 17
    assert capi.Grid.linear(5).regular(5).size == 1471
    assert capi.Grid.linear(8).regular(6).size == 31745
    assert capi.Grid.modlinear(3).regular(4).size == 111


def test_regular_twice_fails():
    g = capi.Grid.linear(2).regular(3)
    with pytest.raises(capi.SGppHIPError) as e:
        g.regular(3)
    assert e.value.status_name == "GENERATION"


def test_refinement_one_to_five():
    # test_GridFactory.cpp:393-407
    g = capi.Grid.linear(2).regular(1)
    assert g.size == 1
    g.refine_surplus(np.array([1.0]), 1)
    assert g.size == 5


def test_regular_order_2d():
    # order fixed by regular_iter: (1,1;1,1) (2,1;1,1) (2,3;1,1) (1,1;2,1) (1,1;2,3)
    lv, ix = capi.Grid.linear(2).regular(2).level_index()
    assert lv.tolist() == [[1, 1], [2, 1], [2, 1], [1, 2], [1, 2]]
    assert ix.tolist() == [[1, 1], [1, 1], [3, 1], [1, 1], [1, 3]]


@pytest.mark.parametrize("basis", [0, 1])
def test_hierarchisation_interpolates(basis):
    # the hierarchical interpolant reproduces f at every grid point (no reference needed)
    g = capi.Grid(basis, 3).regular(4)
    c = g.coordinates()
    f = workloads.gaussian_bump(c)
    a = g.hierarchise(f)
    lv, ix = g.level_index()
    back = oracle.mult(basis, c, lv, ix, a)
    assert np.max(np.abs(back - f)) < 1e-13


@pytest.mark.parametrize("basis", [0, 1])
def test_incremental_equals_full(basis):
    g = capi.Grid(basis, 4).regular(3)
    vals = workloads.gaussian_bump(g.coordinates())
    a = g.hierarchise(vals, incremental=True)
    for _ in range(4):
        g.refine_surplus(a, 30)
        vals = workloads.gaussian_bump(g.coordinates())
        a = g.hierarchise(vals, incremental=True)
        full = g.hierarchise(vals)
        assert (a.view(np.uint64) == full.view(np.uint64)).all()


def test_serialize_roundtrip():
    g = capi.Grid.modlinear(3).regular(3)
    g.refine_surplus(np.linspace(1, 2, g.size), 5)
    text = g.serialize()
    assert text.startswith("modlinear")
    h = capi.Grid.unserialize(text)
    assert h.size == g.size
    for a, b in zip(g.level_index(), h.level_index()):
        assert (a == b).all()
    assert (g.leaf() == h.leaf()).all()
    with pytest.raises(capi.SGppHIPError):
        capi.Grid.unserialize("polynomial\n9 1 1\n")


def test_bad_arguments():
    with pytest.raises(capi.SGppHIPError) as e:
        capi.Grid(capi.BASIS_LINEAR, 0)
    assert e.value.status_name == "INVALID_ARGUMENT"
    with pytest.raises(capi.SGppHIPError) as e:
        capi.Grid(7, 2)
    assert e.value.status_name == "FACTORY"
    g = capi.Grid.linear(2).regular(2)
    with pytest.raises(capi.SGppHIPError) as e:
        g.refine_surplus(np.ones(3), 1)
    assert e.value.status_name == "INVALID_ARGUMENT"


def test_refinement_stops_at_level_31():
    """Refining the left-most point of a 1-D grid again and again: the
    left-most point reaches level 31 after 30 steps; the next step would need
    level 32 and is refused without touching the grid."""
    g = capi.Grid.linear(1).regular(1)
    for step in range(30):
        lv, ix = g.level_index()
        left = int(np.argmin(ix[:, 0] / 2.0 ** lv[:, 0]))
        a = np.zeros(g.size)
        a[left] = 1.0
        g.refine_surplus(a, 1)
    lv, ix = g.level_index()
    assert lv.max() == 31 and g.size == 61
    left = int(np.argmin(ix[:, 0] / 2.0 ** lv[:, 0]))
    assert lv[left, 0] == 31 and ix[left, 0] == 1
    a = np.zeros(g.size)
    a[left] = 1.0
    with pytest.raises(capi.SGppHIPError) as e:
        g.refine_surplus(a, 1)
    assert e.value.status_name == "GENERATION"
    assert g.size == 61


@ref_only
@pytest.mark.parametrize("basis", [0, 1])
@pytest.mark.parametrize("d,L", [(1, 5), (2, 4), (5, 5), (10, 4), (16, 3)])
def test_regular_matches_reference(basis, d, L):
    g = capi.Grid(basis, d).regular(L)
    r = oracle.RefGrid(basis, d).regular(L)
    lv, ix = g.level_index()
    rlv, rix, rlf = r.level_index(with_leaf=True)
    assert (lv == rlv).all() and (ix == rix).all() and (g.leaf() == rlf).all()
    f = workloads.gaussian_bump(g.coordinates())
    a, ra = g.hierarchise(f), r.hierarchise(f)
    assert (a.view(np.uint64) == ra.view(np.uint64)).all()


@ref_only
@pytest.mark.parametrize("basis", [0, 1])
def test_adaptive_matches_reference(basis):
    d = 6
    g, r = capi.Grid(basis, d).regular(3), oracle.RefGrid(basis, d).regular(3)
    for step in range(10):
        f = workloads.cosine_product(g.coordinates())
        a, ra = g.hierarchise(f, incremental=True), r.hierarchise(f)
        assert (a.view(np.uint64) == ra.view(np.uint64)).all(), step
        g.refine_surplus(a, 40)
        r.refine_surplus(ra, 40)
        lv, ix = g.level_index()
        rlv, rix, rlf = r.level_index(with_leaf=True)
        assert g.size == r.size and (lv == rlv).all() and (ix == rix).all(), step
        assert (g.leaf() == rlf).all(), step


@ref_only
def test_modlinear_d10_refinement_matches_reference():
    # the refinement loop of the ModLinear d=10 streaming workload, cut at 20,000 points
    g, _, steps = workloads.adaptive_grid(capi.BASIS_MODLINEAR, 10, 4, workloads.gaussian_bump,
                                          200, 20000)
    r = oracle.RefGrid(capi.BASIS_MODLINEAR, 10).regular(4)
    for _ in range(steps):
        c = (lambda li: li[1] / (2.0 ** li[0]))(r.level_index())
        r.refine_surplus(r.hierarchise(workloads.gaussian_bump(c)), 200)
    lv, ix = g.level_index()
    rlv, rix = r.level_index()
    assert g.size == r.size
    assert (lv == rlv).all() and (ix == rix).all()


def grid_digest(g):
    """SHA-256 of the level then index arrays (uint32 little-endian, N x d in
    sequence order), as tests/golden/make_grid_hashes.py writes them for the
    reference's own refinement loop."""
    import hashlib
    lv, ix = g.level_index()
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(lv, dtype="<u4").tobytes())
    h.update(np.ascontiguousarray(ix, dtype="<u4").tobytes())
    return h.hexdigest()


def test_modlinear_d10_benchmark_grid_matches_reference_hash():
    """The full 100,015-point ModLinear d=10 grid of the benchmark, point for
    point (order included) the grid the reference's own Hierarchisation +
    SurplusRefinementFunctor loop builds (fixture from oracle/_ref)."""
    import json
    with open(os.path.join(GOLDEN, "grid_hashes.json")) as f:
        ref = json.load(f)["modlinear_d10_l4_gaussian_bump_100k"]
    g, _, steps = workloads.adaptive_grid(capi.BASIS_MODLINEAR, 10, 4, workloads.gaussian_bump,
                                          200, 100000)
    assert (g.size, steps) == (ref["points"], ref["refinement_steps"])
    assert grid_digest(g) == ref["sha256"]
