"""Write tests/golden/regression_alpha.npz: the reference's own FP64 CPU
solve of (B^T B + lambda*M*I) alpha = B^T y for the Linear d=8 level-6
regression workload (500,000 points, lambda = 1e-4), with the reference's
ConjugateGradients to a 1e-10 relative residual on its SubspaceLinear operator
(both compiled from the reference sources into oracle/_ref). Run once in the
container that holds the reference; the output is committed."""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402
from sgpp_amd import workloads  # noqa: E402

if __name__ == "__main__":
    if oracle.ref() is None:
        sys.exit("build oracle/_ref first (make -C oracle ref)")
    c = workloads.config_regression()
    lv, ix = c["grid"].level_index()
    g = oracle.RefGrid.from_arrays(0, lv, ix)
    op = oracle.RefOp(g, oracle.KIND_SUBSPACE, c["X"])
    t = time.time()
    alpha, iters, resid = op.cg_solve(c["y"], c["lam"], 10000, 1e-10)
    secs = time.time() - t
    np.savez_compressed(os.path.join(HERE, "regression_alpha.npz"), alpha=alpha,
                        iterations=iters, residuum=resid, seconds=secs,
                        threads=oracle.num_threads())
    print("iterations", iters, "residuum", resid, "seconds", secs)
