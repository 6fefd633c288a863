"""Writes the golden fixtures of the distance path into tests/golden/*.npz.

The gputools reference ships no test vectors (it has no tests; the Rd examples
at man/gpuDist.Rd:33-41 are unseeded and print no expected values), and R is
not installed here, so the expected values come from an independent FP64
implementation, scipy.spatial.distance.pdist (scipy 1.15.3), on inputs where
its definitions coincide with R's stats::dist:

  euclidean  -> pdist 'euclidean'
  maximum    -> pdist 'chebyshev'
  manhattan  -> pdist 'cityblock'
  canberra   -> pdist 'canberra' on strictly positive inputs (scipy divides by
                |u|+|v|, R by |u+v|; equal when u, v > 0)
  binary     -> pdist 'jaccard' on the boolean matrix x != 0 (scipy and R both
                give 0 when no bit is on in either row)
  minkowski  -> pdist 'minkowski' with the same p

pdist's condensed order (i < j, row-major upper triangle) is the same sequence
as R's 'dist' order (lower triangle, column-major).

Run from the repo root:  python tests/golden/make_golden.py
"""
import os

import numpy as np
from scipy.spatial.distance import pdist

HERE = os.path.dirname(os.path.abspath(__file__))

SCIPY_NAME = {"euclidean": "euclidean", "maximum": "chebyshev", "manhattan": "cityblock",
              "canberra": "canberra", "binary": "jaccard", "minkowski": "minkowski"}


def expected(x, method, p):
    xd = x.astype(np.float64)
    if method == "binary":
        return pdist(xd != 0, "jaccard")
    if method == "minkowski":
        return pdist(xd, "minkowski", p=p)
    return pdist(xd, SCIPY_NAME[method])


def case(name, x, method, p=2.0):
    x = np.ascontiguousarray(x, dtype=np.float32)
    np.savez(os.path.join(HERE, name + ".npz"), x=x, method=np.array(method),
             p=np.array(p, dtype=np.float64), expected=expected(x, method, p))


def main():
    rng = np.random.default_rng(20240601)
    # cfg1 shape at reduced n: U[0,1), d = 64
    case("euclidean_uniform_60x64", rng.random((60, 64)), "euclidean")
    g = rng.standard_normal((37, 19))
    case("euclidean_normal_37x19", g, "euclidean")
    case("maximum_normal_37x19", g, "maximum")
    case("manhattan_normal_37x19", g, "manhattan")
    case("canberra_shifted_37x19", np.abs(g) + 0.1, "canberra")
    case("minkowski3_normal_37x19", g, "minkowski", 3.0)
    case("minkowski1p5_normal_37x19", g, "minkowski", 1.5)
    case("binary_bernoulli_41x77", (rng.random((41, 77)) < 0.1), "binary")
    gz = rng.standard_normal((33, 45))
    gz[rng.random(gz.shape) < 0.6] = 0.0
    case("binary_sparse_normal_33x45", gz, "binary")
    # ragged tile edges: n and k not multiples of any tile size
    h = rng.standard_normal((131, 70))
    for m in ("euclidean", "maximum", "manhattan"):
        case(m + "_normal_131x70", h, m)
    case("canberra_shifted_131x70", np.abs(h) + 0.1, "canberra")
    case("minkowski3_normal_131x70", h, "minkowski", 3.0)
    case("binary_bernoulli_131x70", rng.random((131, 70)) < 0.3, "binary")
    # degenerate sizes
    case("euclidean_n2_k1", rng.standard_normal((2, 1)), "euclidean")
    case("manhattan_n1_k5", rng.standard_normal((1, 5)), "manhattan")


if __name__ == "__main__":
    main()
