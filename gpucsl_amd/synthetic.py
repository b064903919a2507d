"""Seeded synthetic datasets for skeleton discovery.

Gaussian: a linear structural equation model on a random Erdos-Renyi DAG.
  A random permutation is the causal order; each forward pair gets an edge
  with probability q = d / (p - 1) (d = expected total degree); coefficients
  are U(0.5, 1) with a random sign; noise is N(0, 1); samples are drawn in
  topological order.
Discrete: the same DAG construction with cardinalities uniform in {2..5};
  every conditional-probability row is Dirichlet(1); ancestral sampling.

All draws use numpy.random.Generator(PCG64(seed)).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class DagData:
    data: np.ndarray          # (n, p) observations x variables
    adjacency: np.ndarray     # (p, p) bool, adjacency[a, b] = edge a -> b
    order: np.ndarray         # causal order (permutation of 0..p-1)
    seed: int


def random_dag(p: int, degree: float, rng: np.random.Generator):
    order = rng.permutation(p)
    q = 0.0 if p < 2 else min(1.0, degree / (p - 1))
    A = np.zeros((p, p), bool)
    # forward pairs in causal order: order[a] -> order[b] for a < b
    upper = rng.random((p, p)) < q
    upper = np.triu(upper, k=1)
    ia, ib = np.nonzero(upper)
    A[order[ia], order[ib]] = True
    return A, order


def gaussian_sem(p: int, n: int, degree: float = 3.0, seed: int = 0,
                 dtype=np.float64) -> DagData:
    rng = np.random.Generator(np.random.PCG64(seed))
    A, order = random_dag(p, degree, rng)
    B = np.where(A, rng.uniform(0.5, 1.0, (p, p)) * rng.choice([-1.0, 1.0], (p, p)), 0.0)
    X = np.empty((n, p), np.float64)
    noise = rng.standard_normal((n, p))
    for k in order:
        parents = np.flatnonzero(A[:, k])
        x = noise[:, k].copy()
        if parents.size:
            x += X[:, parents] @ B[parents, k]
        X[:, k] = x
    return DagData(X.astype(dtype, copy=False), A, order, seed)


def discrete_bn(p: int, n: int, degree: float = 2.0, seed: int = 0,
                card_low: int = 2, card_high: int = 5) -> DagData:
    """Returns uint8 codes (n, p); column k takes values in 0..card_k-1."""
    rng = np.random.Generator(np.random.PCG64(seed))
    A, order = random_dag(p, degree, rng)
    card = rng.integers(card_low, card_high + 1, size=p)
    X = np.zeros((n, p), np.uint8)
    for k in order:
        parents = np.flatnonzero(A[:, k])
        cfg = np.zeros(n, np.int64)
        radix = 1
        for pa in parents:
            cfg += X[:, pa].astype(np.int64) * radix
            radix *= int(card[pa])
        cpt = rng.dirichlet(np.ones(int(card[k])), size=radix)   # (configs, card_k)
        cdf = np.cumsum(cpt, axis=1)
        u = rng.random(n)
        vals = (u[:, None] > cdf[cfg]).sum(axis=1)
        X[:, k] = np.minimum(vals, card[k] - 1).astype(np.uint8)
    return DagData(X, A, order, seed)


def bootstrap(data: np.ndarray, seed: int) -> np.ndarray:
    rng = np.random.Generator(np.random.PCG64(seed))
    rows = rng.integers(0, data.shape[0], size=data.shape[0])
    return data[rows]
