"""Writes tests/golden/reference_vectors.json.

Known-answer vectors of the reference's own tests, transcribed as numbers
(the reference is not imported or run).  Where the reference test derives its
expectation from numpy/pandas at run time, the expectation is recomputed here
with numpy from the same seeded inputs.

Sources (gpucsl repository):
  tests/cuda/helper_test.cc:10-17          binomial_coefficient
  tests/cuda/helper_test.cc:25-31          pchisq (EXPECT_FLOAT_EQ)
  tests/test_cuda_set_helpers.py:15-42     5C3 lexicographic unranking
  tests/test_cuda_graph_helpers.py:15-69   compact on a 6x6 random and full graph
  tests/test_cuda_gaussian_helpers.py:76-127  inverse / Cholesky / pinv inputs
  tests/test_cuda_discrete_helpers.py:20-172  contingency, marginals, chi^2
Run:  python tests/golden/make_golden.py
"""
import itertools
import json
from pathlib import Path

import numpy as np

OUT = Path(__file__).resolve().parent / "reference_vectors.json"


def main():
    g = {}
    g["binomial"] = [[2, 3, 0], [3, 2, 3], [5, 2, 10], [20, 10, 184756], [30, 20, 30045015],
                     [30, 30, 1]]
    g["pchisq"] = [[3.0, 2, 0.22313016014842982], [20.0, 10, 0.02925268807696113],
                   [0.0, 1, 1.0], [10.0, 10, 0.4404932850652121],
                   [0.2, 3, 0.9775892977616494]]
    g["unrank_5c3"] = [[1, 2, 3], [1, 2, 4], [1, 2, 5], [1, 3, 4], [1, 3, 5], [1, 4, 5],
                       [2, 3, 4], [2, 3, 5], [2, 4, 5], [3, 4, 5]]
    g["compact_random"] = {
        "G": [[0, 1, 1, 0, 0, 1], [1, 0, 1, 0, 1, 0], [0, 0, 0, 0, 0, 0],
              [1, 1, 1, 0, 1, 1], [1, 0, 1, 0, 0, 1], [0, 1, 1, 1, 0, 0]],
        "Gc": [[3, 1, 2, 5, 0, 0], [3, 0, 2, 4, 0, 0], [0, 0, 0, 0, 0, 0],
               [5, 0, 1, 2, 4, 5], [3, 0, 2, 5, 0, 0], [3, 1, 2, 3, 0, 0]],
    }
    g["compact_full"] = {
        "G": np.ones((6, 6), int).tolist(),
        "Gc": [[5, 1, 2, 3, 4, 5], [5, 0, 2, 3, 4, 5], [5, 0, 1, 3, 4, 5],
               [5, 0, 1, 2, 4, 5], [5, 0, 1, 2, 3, 5], [5, 0, 1, 2, 3, 4]],
    }
    inv_in = np.array([[5, 3, 4], [3, 4, 5], [6, 3, 2]], float)
    chol_in = np.array([[4.2, 12.3, -16.4], [12.3, 37.3, -43.3], [-16.4, -43.5, 98.6]], float)
    pinv_in = np.array([[1, 2, 3], [5, 2, 2], [1, 6, 1]], float)
    g["inverse3"] = {"A": inv_in.tolist(), "expected": np.linalg.inv(inv_in).tolist()}
    g["cholesky3"] = {"A": chol_in.tolist(), "expected": np.linalg.cholesky(chol_in).tolist()}
    g["m2_inv3"] = {"A": pinv_in.tolist(), "expected": np.linalg.pinv(pinv_in).tolist()}

    a = np.array([0, 1, 2, 3, 4, 5, 1, 2, 3, 4, 5, 6])
    b = np.array([0, 1, 1, 1, 1, 0, 0, 1, 0, 0, 1, 0])
    ct = np.zeros((7, 2), int)
    np.add.at(ct, (a, b), 1)
    g["contingency_level0"] = {"a": a.tolist(), "b": b.tolist(), "expected": ct.tolist()}

    np.random.seed(1)
    data = np.random.randint(0, 5, size=(100, 4))
    counts = np.zeros((5, 5, 5, 5), int)
    for row in data:
        counts[row[1], row[0], row[2], row[3]] += 1
    g["contingency_level2"] = {"data": data.tolist(), "vi": 2, "vj": 3, "sepset": [0, 1],
                               "dims": [5, 5, 5, 5], "expected": counts.flatten().tolist()}

    np.random.seed(1)
    N = np.random.randint(1, 100, size=(5, 5, 5))
    Ns = np.zeros(5)
    Nis = np.zeros((5, 5))
    Njs = np.zeros((5, 5))
    for s, vi, vj in itertools.product(range(5), repeat=3):
        v = N[s, vi, vj]
        Ns[s] += v
        Nis[s, vi] += v
        Njs[s, vj] += v
    chi2 = 0.0
    for s, vi, vj in itertools.product(range(5), repeat=3):
        o = N[s, vi, vj]
        e = Nis[s, vi] * Njs[s, vj] / Ns[s]
        chi2 += (o - e) * (o - e) / e
    g["marginals_chi2"] = {"N": N.flatten().tolist(), "Nis": Nis.astype(int).flatten().tolist(),
                           "Njs": Njs.astype(int).flatten().tolist(), "Ns": Ns.astype(int).tolist(),
                           "dim_s": 5, "di": 5, "dj": 5, "chi2": chi2, "tol": 1e-10}
    OUT.write_text(json.dumps(g, indent=1))
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
