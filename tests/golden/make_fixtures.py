"""Write the golden fixtures of the parity tests (run once, in the container
that holds the reference at /root/reference; the outputs are committed).

  bt_bbt.json     For every Linear ("hut") and ModLinear ("ausgeklappt")
                  golden matrix of datadriven/datasets/BT_BBT: the data points
                  and the order-independent quantities the reference's
                  testBT.cpp / testBBT.cpp compare (BT row sums in data order,
                  sorted BT column sums; sorted BBT diagonal, row and column
                  sums).
  friedman.npz    The reference's differential-test datasets
                  (datadriven/datasets/friedman/friedman2_4d_10000.arff.gz,
                  friedman1_10d_2000.arff.gz; the first 2000 rows of each),
                  the grid of compareToReference (regular level 5 + one surplus
                  refinement of 80 points; the refinement indicator is seeded
                  here, where the reference uses std::random_device) and the
                  outputs of the reference's DEFAULT operator (the compiled
                  reference of oracle/_ref) for alpha_i = i and y_i = i + 1
                  (test_datadrivenCommon.cpp:205-341).
"""
import gzip
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from sgpp_amd import workloads  # noqa: E402


def read_arff(path, with_class):
    rows = []
    started = False
    with gzip.open(path, "rt") as f:
        for line in f:
            line = line.strip()
            if not line:
                continue
            if not started:
                if line.upper().startswith("@DATA"):
                    started = True
                continue
            rows.append([float(v) for v in line.split(",")])
    a = np.array(rows)
    return (a[:, :-1], a[:, -1]) if with_class else (a, None)


def read_matrix(path):
    with gzip.open(path, "rt") as f:
        return np.array([[float(v) for v in line.split()] for line in f if line.strip()])


BT_CASES = [
    # (basis, name, dim, level, data file, nopsgrid)
    (0, "hut", 1, 3, "data_dim_1_nops_8_float.arff.gz", 7),
    (0, "hut", 1, 5, "data_dim_1_nops_8_float.arff.gz", 31),
    (0, "hut", 3, 3, "data_dim_3_nops_512_float.arff.gz", 31),
    (0, "hut", 3, 4, "data_dim_3_nops_512_float.arff.gz", 111),
    (1, "ausgeklappt", 1, 3, "data_dim_1_nops_8_float.arff.gz", 7),
    (1, "ausgeklappt", 1, 5, "data_dim_1_nops_8_float.arff.gz", 31),
    (1, "ausgeklappt", 3, 3, "data_dim_3_nops_512_float.arff.gz", 31),
    (1, "ausgeklappt", 3, 4, "data_dim_3_nops_512_float.arff.gz", 111),
]


def bt_bbt():
    base = os.path.join(REF, "datadriven", "datasets", "BT_BBT")
    out = []
    for basis, name, dim, level, dfile, nops in BT_CASES:
        X, _ = read_arff(os.path.join(base, dfile), False)
        bt = read_matrix(os.path.join(base, "BT_phi_li_%s_dim_%d_nopsgrid_%d_float.dat.gz"
                                      % (name, dim, nops)))
        bbt = read_matrix(os.path.join(base, "BBT_phi_li_%s_dim_%d_nopsgrid_%d_float.dat.gz"
                                       % (name, dim, nops)))
        assert bt.shape == (X.shape[0], nops) and bbt.shape == (nops, nops)
        out.append(dict(basis=basis, dim=dim, level=level, nops=nops, data=X.tolist(),
                        bt_row_sums=bt.sum(1).tolist(),
                        bt_col_sums_sorted=sorted(bt.sum(0).tolist()),
                        bbt_diag_sorted=sorted(np.diag(bbt).tolist()),
                        bbt_row_sums_sorted=sorted(bbt.sum(1).tolist()),
                        bbt_col_sums_sorted=sorted(bbt.sum(0).tolist()),
                        source="datadriven/datasets/BT_BBT/BT_phi_li_%s_dim_%d_nopsgrid_%d_float"
                               % (name, dim, nops)))
    with open(os.path.join(HERE, "bt_bbt.json"), "w") as f:
        json.dump(out, f)


def friedman():
    base = os.path.join(REF, "datadriven", "datasets", "friedman")
    arrays = {}
    for tag, fname in [("f2", "friedman2_4d_10000.arff.gz"), ("f1", "friedman1_10d_2000.arff.gz")]:
        X, _ = read_arff(os.path.join(base, fname), True)
        X = X[:2000]
        arrays[tag + "_X"] = X
        for basis in (0, 1):
            g = oracle.RefGrid(basis, X.shape[1]).regular(5)
            refine = 1.0 + 99.0 * workloads.uniform(workloads.stream(7, basis), g.size)
            g.refine_surplus(refine, 80, 0.0)
            lv, ix = g.level_index()
            alpha = np.arange(g.size, dtype=np.float64)
            y = np.arange(1, X.shape[0] + 1, dtype=np.float64)
            op = oracle.RefOp(g, oracle.KIND_DEFAULT, X)
            key = "%s_b%d" % (tag, basis)
            arrays[key + "_level"] = lv.astype(np.uint8)
            arrays[key + "_index"] = ix.astype(np.uint8)
            arrays[key + "_mult"] = op.mult(alpha)
            arrays[key + "_multT"] = op.mult_transpose(y)
    np.savez_compressed(os.path.join(HERE, "friedman.npz"), **arrays)


if __name__ == "__main__":
    if oracle.ref() is None:
        sys.exit("build oracle/_ref first (make -C oracle ref)")
    bt_bbt()
    friedman()
    print("fixtures written to", HERE)
