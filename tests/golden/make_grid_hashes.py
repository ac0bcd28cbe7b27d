"""Write tests/golden/grid_hashes.json: SHA-256 of the level and index arrays
(uint32 little-endian, N x d in sequence order, level array then index
array) of the two adaptive grids of the benchmarks, built by the REFERENCE's
own loop: Grid::createModLinearGrid / createLinearGrid, regular start grid,
then repeatedly OperationHierarchisation of f at the grid points and
HashRefinement with SurplusRefinementFunctor(alpha, 200) until the grid has
at least the target size (compiled from the reference sources into
oracle/_ref). f is evaluated by sgpp_amd.workloads' numpy functions at the
points' coordinates i / 2^l, as the benchmarks' generator does.

  ModLinear d=10 from level 4, f = exp(-sum (x_j - 0.5)^2 / 0.05), >= 100,000 points
  Linear d=20 from level 4, f = prod 4 x_j (1 - x_j) (1 + 0.5 cos(pi x_j)), >= 1,000,000

Run once in the container that holds the reference; the output is committed."""
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402
from sgpp_amd import workloads  # noqa: E402

CASES = [
    ("modlinear_d10_l4_gaussian_bump_100k", oracle.BASIS_MODLINEAR, 10, 4,
     workloads.gaussian_bump, 200, 100000),
    ("linear_d20_l4_damped_cosine_product_1m", oracle.BASIS_LINEAR, 20, 4,
     workloads.damped_cosine_product, 200, 1000000),
]


def digest(lv, ix):
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(lv, dtype="<u4").tobytes())
    h.update(np.ascontiguousarray(ix, dtype="<u4").tobytes())
    return h.hexdigest()


def reference_loop(basis, dim, start_level, f, per_step, target):
    g = oracle.RefGrid(basis, dim)
    g.regular(start_level)
    lv, ix = g.level_index()
    vals = f(ix / (2.0 ** lv))
    alpha = g.hierarchise(vals)
    steps = 0
    while g.size < target:
        n0 = g.size
        g.refine_surplus(alpha, per_step)
        if g.size == n0:
            break
        lv, ix = g.level_index()
        vals = np.concatenate([vals, f(ix[n0:] / (2.0 ** lv[n0:]))])
        alpha = g.hierarchise(vals)
        steps += 1
        if steps % 10 == 0:
            print("  step %d: %d points" % (steps, g.size), flush=True)
    lv, ix = g.level_index()
    return lv, ix, steps


if __name__ == "__main__":
    if oracle.ref() is None:
        sys.exit("build oracle/_ref first (make -C oracle ref)")
    out = {}
    for name, basis, dim, start, f, per_step, target in CASES:
        t = time.time()
        lv, ix, steps = reference_loop(basis, dim, start, f, per_step, target)
        out[name] = {"points": int(lv.shape[0]), "dim": dim, "refinement_steps": steps,
                     "sha256": digest(lv, ix), "seconds": round(time.time() - t, 1)}
        print(name, out[name], flush=True)
    with open(os.path.join(HERE, "grid_hashes.json"), "w") as fh:
        json.dump(out, fh, indent=1)
