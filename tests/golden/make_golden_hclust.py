"""Writes the golden fixtures of the clustering stage into tests/golden/hclust/.

The gputools reference has no test vectors for gpuHclust / gpuDistClust, and
R is not installed here.  The expected merges and heights come from an
independent implementation, scipy.cluster.hierarchy.linkage (scipy 1.15.3),
for the methods whose Lance-Williams updates it shares with the reference
(inst/cuda/hcluster.cu):

  single, complete, average     -> linkage 'single', 'complete', 'average'
  mcquitty, wpgma               -> linkage 'weighted'
  median, centroid              -> linkage 'median', 'centroid' on sqrt(d):
                                   scipy applies these updates to squared
                                   distances, the reference to d itself, so
                                   the expected heights are scipy's squared.

Each fixture holds the FP32 distance matrix d (full n x n, from FP64 pdist
of seeded N(0,1) points, rounded to FP32), and per method the merge matrix in
R's hclust convention (1-based, negative = singleton, singletons first, then
clusters in increasing order; two singletons by increasing index) and the
FP64 heights.  Inputs are random, so no two candidate distances tie.

Run from the repo root:  python tests/golden/make_golden_hclust.py
"""
import os

import numpy as np
from scipy.cluster.hierarchy import linkage
from scipy.spatial.distance import pdist, squareform

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hclust")

METHODS = {"single": ("single", False), "complete": ("complete", False),
           "average": ("average", False), "mcquitty": ("weighted", False),
           "wpgma": ("weighted", False), "median": ("median", True),
           "centroid": ("centroid", True)}


def r_merge(Z, n):
    """scipy's Z (cluster n + s formed at step s) in R's merge convention."""
    rows = []
    for a, b in Z[:, :2].astype(np.int64):
        ea = -(a + 1) if a < n else a - n + 1
        eb = -(b + 1) if b < n else b - n + 1
        if ea < 0 and eb < 0:
            ea, eb = max(ea, eb), min(ea, eb)
        elif ea > 0 > eb:
            ea, eb = eb, ea
        elif ea > 0 and eb > 0 and ea > eb:
            ea, eb = eb, ea
        rows.append((ea, eb))
    return np.array(rows, dtype=np.int32).reshape(-1, 2)


def case(name, x):
    d32 = squareform(pdist(x.astype(np.float64))).astype(np.float32)
    n = d32.shape[0]
    cond = squareform(d32.astype(np.float64), checks=False)
    out = {"d": d32}
    for m, (scipy_m, squared) in METHODS.items():
        Z = linkage(np.sqrt(cond) if squared else cond, scipy_m)
        h = Z[:, 2] ** 2 if squared else Z[:, 2]
        out["merge_" + m] = r_merge(Z, n)
        out["height_" + m] = h.astype(np.float64)
    np.savez(os.path.join(HERE, name + ".npz"), **out)


def main():
    os.makedirs(HERE, exist_ok=True)
    rng = np.random.default_rng(20240917)
    case("normal_40x5", rng.standard_normal((40, 5)))
    case("normal_12x3", rng.standard_normal((12, 3)))
    case("normal_97x16", rng.standard_normal((97, 16)))
    case("normal_3x2", rng.standard_normal((3, 2)))
    case("normal_2x4", rng.standard_normal((2, 4)))


if __name__ == "__main__":
    main()
