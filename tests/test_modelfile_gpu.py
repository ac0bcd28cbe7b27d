"""A model trained by the reference, read from its own files, served and
evaluated through the operator on the GPU: the results are the reference
test's expected values (test_Classifier.py:89, 7 places) and their signs the
classes the reference wrote (classifier.apply.arff.out.arff). FP32 products
are held to 1e-6 absolute on these O(0.1) values."""
import os

import numpy as np
import pytest

from sgpp_amd import modelfile

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(__file__), "golden", "reference_model")
APPLY_EXPECTED = np.array([0.253400605292, -0.25507958758, 0.0530555506998])


def load():
    grid, alpha = modelfile.load_model(os.path.join(HERE, "classifier_test.a0.grid"),
                                       os.path.join(HERE, "classifier_test.a0.alpha.arff"))
    X, _ = modelfile.read_arff(os.path.join(HERE, "classifier.apply.arff"))
    _, cls = modelfile.read_arff(os.path.join(HERE, "classifier.apply.arff.out.arff"))
    return grid, alpha, X, cls


@pytest.mark.parametrize("prec,tol", [(0, 5e-8), (1, 1e-6)])
def test_served_reference_model(gpu, prec, tol):
    capi = gpu
    grid, alpha, X, cls = load()
    lv, ix = grid.level_index()
    srv = capi.Server(precision=prec, memory_budget=1 << 30, max_batch_queries=1 << 16,
                      max_items=1 << 12, max_partial_doubles=1 << 20)
    mid = srv.admit_model(grid.basis, lv, ix, alpha)
    rid, n = srv.submit(mid, X)
    st, r, _ = srv.poll(rid, n)
    assert st == capi.REQ_DONE
    assert np.abs(r - APPLY_EXPECTED).max() < tol
    assert np.array_equal(np.sign(r), cls)
    srv.close()


@pytest.mark.parametrize("prec,tol", [(0, 5e-8), (1, 1e-6)])
def test_operator_on_reference_model(gpu, prec, tol):
    capi = gpu
    grid, alpha, X, cls = load()
    op = capi.MultiEval(grid.basis, X, precision=prec, grid=grid).prepare()
    r = op.mult(alpha)
    assert np.abs(r - APPLY_EXPECTED).max() < tol
    assert np.array_equal(np.sign(r), cls)
