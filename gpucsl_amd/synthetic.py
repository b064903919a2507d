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


# ---------------------------------------------------------------------------
# Device generators for large request streams.  The DAG, coefficients and
# conditional-probability tables come from numpy (as above); the n x p random
# draws come from a seeded torch generator on the GPU, so a 2000 x 50000
# dataset takes milliseconds instead of seconds.  The values differ from the
# numpy generators' for the same seed.
# ---------------------------------------------------------------------------
def gaussian_sem_device(p: int, n: int, degree: float = 3.0, seed: int = 0, device="cuda"):
    """Float32 (n, p) tensor on the device: X = E (I - B)^-1 for the linear SEM
    on a random ER DAG; (I - B) is unit upper triangular in causal order."""
    import torch

    rng = np.random.Generator(np.random.PCG64(seed))
    A, order = random_dag(p, degree, rng)
    B = np.where(A, rng.uniform(0.5, 1.0, (p, p)) * rng.choice([-1.0, 1.0], (p, p)), 0.0)
    Bo = B[np.ix_(order, order)]                    # strictly upper triangular in causal order
    g = torch.Generator(device=device)
    g.manual_seed(int(seed))
    E = torch.randn((n, p), generator=g, device=device, dtype=torch.float64)
    I = torch.eye(p, dtype=torch.float64, device=device)
    Minv = torch.linalg.solve_triangular(I - torch.from_numpy(Bo).to(device), I, upper=True)   # p x p
    Xo = E @ Minv
    X = torch.empty_like(Xo)
    X[:, torch.from_numpy(order).to(device)] = Xo
    return X.to(torch.float32)


def discrete_bn_device(p: int, n: int, degree: float = 2.0, seed: int = 0, card_low: int = 2, card_high: int = 5,
                       device="cuda", row_align: int = 64):
    """Ancestral sample of a random discrete BN on the device.  Returns
    (codes (p, stride) uint8 with one row per variable, stride = n rounded up
    to row_align, zero padded; dims (p,) uint8)."""
    import torch

    rng = np.random.Generator(np.random.PCG64(seed))
    A, order = random_dag(p, degree, rng)
    card = rng.integers(card_low, card_high + 1, size=p)
    stride = (n + row_align - 1) // row_align * row_align
    codes = torch.zeros((p, stride), dtype=torch.uint8, device=device)
    g = torch.Generator(device=device)
    g.manual_seed(int(seed))
    for k in order:
        parents = np.flatnonzero(A[:, k])
        cfg = torch.zeros(n, dtype=torch.int64, device=device)
        radix = 1
        for pa in parents:
            cfg += codes[pa, :n].to(torch.int64) * radix
            radix *= int(card[pa])
        cdf = torch.from_numpy(np.cumsum(rng.dirichlet(np.ones(int(card[k])), size=radix), axis=1)).to(device)
        u = torch.rand(n, generator=g, device=device, dtype=torch.float64)
        vals = (u[:, None] > cdf[cfg]).sum(dim=1).clamp_(max=int(card[k]) - 1)
        codes[k, :n] = vals.to(torch.uint8)
    # ordinal encoding (codes of the categories that occur, in sorted order),
    # as the discrete skeleton's host preparation does
    body = codes[:, :n].to(torch.int64)
    bins = 8
    seen = torch.zeros(p * bins, dtype=torch.int64, device=device)
    seen.index_add_(0, (body + torch.arange(p, device=device)[:, None] * bins).reshape(-1),
                    torch.ones(p * n, dtype=torch.int64, device=device))
    present = (seen.view(p, bins) > 0).to(torch.int64)
    remap = present.cumsum(dim=1) - 1
    codes[:, :n] = torch.gather(remap, 1, body).to(torch.uint8)
    dims = present.sum(dim=1).to(torch.uint8)
    return codes, dims


def batched_workload(n_boot: int, n_sets: int):
    """Bootstrap resamples of one p=200, n=5000 dataset (degree 3, seed 300;
    resample b draws its rows with seed 300+b) and uneven datasets k with
    p ~ U[50, 500], n ~ U[1000, 20000] drawn with seed 400+k (degree 3, seed
    400+k).  Returns the base data, the row indices of every resample and the
    uneven datasets (float32)."""
    base = gaussian_sem(200, 5000, 3.0, seed=300).data
    rows = [np.random.Generator(np.random.PCG64(300 + b)).integers(0, 5000, size=5000) for b in range(n_boot)]
    sets = []
    for k in range(n_sets):
        rng = np.random.Generator(np.random.PCG64(400 + k))
        p, n = int(rng.integers(50, 501)), int(rng.integers(1000, 20001))
        sets.append(gaussian_sem(p, n, 3.0, seed=400 + k).data.astype(np.float32))
    return base.astype(np.float32), rows, sets
