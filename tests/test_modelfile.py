"""Trained-model files: the reference's own serialized grid and alpha file,
grid serialization versions 1-9, and the ARFF reader.

The fixtures under tests/golden/reference_model are data files the
reference's tests hold (datadriven/tests/tbin/tlearner/): a Linear d=2
level-2 model trained by the reference (classifier_test.a0.grid, format
version 2, and classifier_test.a0.alpha.arff), the points it is applied to
(datasets/classifier.apply.arff) and the classes written for them
(datasets/classifier.apply.arff.out.arff). The expected evaluations are the
reference test's own numbers (test_Classifier.py:89, checked to 7 places)."""
import os

import numpy as np
import pytest

import oracle
from sgpp_amd import capi, modelfile

HERE = os.path.join(os.path.dirname(__file__), "golden", "reference_model")
GRID = os.path.join(HERE, "classifier_test.a0.grid")
ALPHA = os.path.join(HERE, "classifier_test.a0.alpha.arff")
APPLY = os.path.join(HERE, "classifier.apply.arff")
APPLY_OUT = os.path.join(HERE, "classifier.apply.arff.out.arff")
TRAIN = os.path.join(HERE, "classifier.train.arff")
# test_Classifier.py:89 (Classifier.applyData on the three apply points)
APPLY_EXPECTED = np.array([0.253400605292, -0.25507958758, 0.0530555506998])


def test_reference_model_loads():
    grid, alpha = modelfile.load_model(GRID, ALPHA)
    assert grid.basis == capi.BASIS_LINEAR and grid.dim == 2 and grid.size == 5
    lv, ix = grid.level_index()
    # the file's points in the file's order (sequence numbers)
    assert lv.tolist() == [[1, 1], [2, 1], [2, 1], [1, 2], [1, 2]]
    assert ix.tolist() == [[1, 1], [1, 1], [3, 1], [1, 1], [1, 3]]
    assert grid.leaf().tolist() == [0, 1, 1, 1, 1]
    assert alpha.tolist() == [-0.03105750236900505640, -0.61865507797281660274,
                              0.64903026441541211700, 0.64903026441541256109,
                              -0.61865507797281604763]


def test_reference_model_evaluates_to_the_reference_results():
    grid, alpha = modelfile.load_model(GRID, ALPHA)
    lv, ix = grid.level_index()
    X, cls = modelfile.read_arff(APPLY)
    assert X.tolist() == [[0.5, 0.1], [0.3, 0.4], [0.9, 0.7]]
    r = oracle.mult(capi.BASIS_LINEAR, X, lv, ix, alpha)
    assert np.abs(r - APPLY_EXPECTED).max() < 5e-8
    # the classes the reference wrote for these points are the signs
    _, out_cls = modelfile.read_arff(APPLY_OUT)
    assert np.array_equal(np.sign(r), out_cls)


def _text(version, lv, ix, leaf, dim):
    """Serialized text of a unit-box grid in the given format version
    (layout of HashGridStorage::serialize and HashGridPoint::serialize)."""
    out = ["linear", "%d %d %d" % (version, dim, len(lv))]
    if version >= 5:
        out.append("0")
    if version >= 3:
        out.append(" ".join("0.0 1.0 0 0" for _ in range(dim)))
    for j in range(len(lv)):
        out.append(str(dim))
        out.append(" ".join("%d %d" % (lv[j, d], ix[j, d]) for d in range(dim)))
        if version >= 2 and version != 4:
            out.append(str(int(leaf[j])))
        if 6 <= version <= 8:
            out.append("0")
    return "\n".join(out) + "\n"


@pytest.mark.parametrize("version", range(1, 10))
def test_unserialize_every_version(version):
    g = capi.Grid.linear(3).regular(3)
    g.refine_surplus(np.linspace(1.0, 2.0, g.size), 5)
    lv, ix = g.level_index()
    leaf = g.leaf()
    h = capi.Grid.unserialize(_text(version, lv, ix, leaf, 3))
    lv2, ix2 = h.level_index()
    assert np.array_equal(lv, lv2) and np.array_equal(ix, ix2)
    # versions 1 and 4 store no leaf flags: they are recomputed from the
    # points' children, which must give the flags refinement maintained
    assert np.array_equal(h.leaf(), leaf)


def test_recomputed_leaf_flags_are_not_copied():
    g = capi.Grid.linear(2).regular(3)
    lv, ix = g.level_index()
    wrong = 1 - g.leaf()
    assert np.array_equal(capi.Grid.unserialize(_text(2, lv, ix, wrong, 2)).leaf(), wrong)
    assert np.array_equal(capi.Grid.unserialize(_text(1, lv, ix, wrong, 2)).leaf(), g.leaf())


def test_unserialize_rejects_unknown_versions():
    g = capi.Grid.linear(2).regular(2)
    lv, ix = g.level_index()
    for bad in (0, 10):
        with pytest.raises(capi.SGppHIPError):
            capi.Grid.unserialize(_text(9, lv, ix, g.leaf(), 2).replace("9 2", "%d 2" % bad, 1))


def test_read_arff_rules():
    text = "\n".join([
        "% a comment",
        "@RELATION r",
        "@ATTRIBUTE a NUMERIC",
        "",
        "@DATA",
        "1.5, 2, 1",
        "3e-1,  -4.25abc, -1",
        "x, .5, 1 % trailing comment drops the line",
        "7,8,2",
    ]) + "\n"
    X, y = modelfile.read_arff(text)
    assert X.tolist() == [[1.5, 2.0], [0.3, -4.25], [7.0, 8.0]]
    assert y.tolist() == [1.0, -1.0, 2.0]
    X, y = modelfile.read_arff(text, selected_targets=[1.0, 2.0])
    assert X.tolist() == [[1.5, 2.0], [7.0, 8.0]] and y.tolist() == [1.0, 2.0]
    X, _ = modelfile.read_arff(text, selected_cols=[1], instance_cutoff=2)
    assert X.tolist() == [[2.0], [-4.25]]
    X, y = modelfile.read_arff(text, has_targets=False)
    assert X.shape == (3, 3) and y is None
    with pytest.raises(modelfile.ARFFError):
        modelfile.read_arff("@DATA\n1,2,3\n4,5\n")
    with pytest.raises(modelfile.ARFFError):
        modelfile.read_arff(text, selected_cols=[2])


def test_read_training_set():
    X, y = modelfile.read_arff(TRAIN)
    assert X.shape == (9, 2)
    assert np.allclose(X[:, 0], np.arange(1, 10) / 10) and np.allclose(X[:, 1], 1 - X[:, 0])
    assert y.tolist() == [-1, -1, -1, 1, -1, 1, 1, -1, 1]


def test_alpha_file_must_match_the_grid():
    with open(ALPHA) as f:
        short = f.read().rsplit("\n", 2)[0] + "\n"
    with pytest.raises(modelfile.ARFFError):
        modelfile.load_model(GRID, short)


def test_alpha_writer_reproduces_the_reference_file(tmp_path):
    _, alpha = modelfile.load_model(GRID, ALPHA)
    out = str(tmp_path / "a.arff")
    modelfile.write_alpha(out, alpha, relation="classifier_test.a0.alpha.arff")
    with open(ALPHA) as f:
        want = f.read()
    with open(out) as f:
        got = f.read()
    # the reference names the relation after its output path
    assert got.split("\n")[1:] == want.split("\n")[1:len(got.split("\n"))]
    assert modelfile.read_alpha(out).tolist() == alpha.tolist()


def test_save_and_load_model(tmp_path):
    g = capi.Grid.modlinear(4).regular(3)
    g.refine_surplus(np.linspace(1.0, 2.0, g.size), 4)
    alpha = np.linspace(-1.0, 1.0, g.size)
    gp, ap = str(tmp_path / "m.grid"), str(tmp_path / "m.alpha.arff")
    modelfile.save_model(gp, ap, g, alpha)
    h, a = modelfile.load_model(gp, ap)
    assert h.basis == capi.BASIS_MODLINEAR
    for x, y in zip(g.level_index(), h.level_index()):
        assert np.array_equal(x, y)
    assert np.array_equal(g.leaf(), h.leaf())
    assert np.abs(a - alpha).max() <= 1e-20
