"""Test infrastructure, not product code: a brute-force numpy mini-oracle.

An independent restatement of PC-stable for small p (<= ~12) that shares no
code with pc_oracle.c.  It uses textbook formulas instead of the reference's
operation order:

* Gaussian: rho_{ij|S} = -P_ij / sqrt(P_ii P_jj) with P = inv(C[{i,j}+S]);
  Fisher z = atanh|rho|; removal when z < thr[l]
  (gpucsl/pc/cuda/gaussian_ci.cu:66-120 for the test, helpers.py:42-50 for thr).
* Discrete: Pearson X^2 from np.bincount tables and scipy.stats.chi2.sf
  (gpucsl/pc/cuda/discrete_ci.cu:37-246).

Subsets are enumerated with itertools.combinations (lexicographic, like the
reference's unranking, set_helpers.cu:25-57).  Decisions closer than `margin`
to the threshold are returned as "ties" so a caller can tell a real mismatch
from rounding.
"""
from __future__ import annotations

import itertools

import numpy as np
from scipy.stats import chi2 as _chi2


def _z(C, i, j, S):
    idx = [i, j] + list(S)
    P = np.linalg.inv(C[np.ix_(idx, idx)])
    rho = -P[0, 1] / np.sqrt(P[0, 0] * P[1, 1])
    return abs(np.arctanh(rho))


def gauss_skeleton(C: np.ndarray, thr: np.ndarray, max_level: int, margin: float = 1e-9):
    p = C.shape[0]
    G = np.ones((p, p), bool)
    np.fill_diagonal(G, False)
    sep = {}
    ties = 0
    for a in range(p):
        for b in range(a + 1, p):
            z = abs(np.arctanh(C[a, b]))
            ties += abs(z - thr[0]) < margin
            if z < thr[0]:
                G[a, b] = G[b, a] = False
                sep[(a, b)] = ()
    for l in range(1, max_level + 1):
        nbrs = [list(np.flatnonzero(G[r])) for r in range(p)]
        if max(len(x) for x in nbrs) - 1 < l:
            break
        removed = []
        for a in range(p):
            for b in nbrs[a]:
                if b <= a:
                    continue
                found = None
                for vi, vj in ((a, b), (b, a)):
                    cand = [v for v in nbrs[vi] if v != vj]
                    for S in itertools.combinations(cand, l):
                        z = _z(C, vi, vj, S)
                        ties += abs(z - thr[l]) < margin
                        if z < thr[l]:
                            found = S
                            break
                    if found is not None:
                        break
                if found is not None:
                    removed.append((a, b, found))
        for a, b, S in removed:
            G[a, b] = G[b, a] = False
            sep[(a, b)] = tuple(S)
    return G.astype(np.uint16), sep, int(ties)


def _pearson(data, di, dj, vi, vj, S, dims):
    g = np.zeros(data.shape[1], np.int64)
    radix = 1
    for s in S:
        g += data[s].astype(np.int64) * radix
        radix *= int(dims[s])
    cells = (g * di + data[vi]) * dj + data[vj]
    N = np.bincount(cells, minlength=radix * di * dj).reshape(radix, di, dj).astype(np.float64)
    Nis = N.sum(axis=2)
    Njs = N.sum(axis=1)
    Ns = N.sum(axis=(1, 2))
    stat = 0.0
    for gg in range(radix):
        if Ns[gg] == 0:
            continue
        E = np.outer(Nis[gg], Njs[gg]) / Ns[gg]
        m = E != 0
        stat += (((N[gg] - E) ** 2)[m] / E[m]).sum()
    df = (di - 1) * (dj - 1) * radix
    return stat, df


def disc_skeleton(data: np.ndarray, dims: np.ndarray, alpha: float, max_level: int,
                  margin: float = 1e-9):
    """data: column-major (p, n) integer codes."""
    p = data.shape[0]
    G = np.ones((p, p), bool)
    np.fill_diagonal(G, False)
    sep = {}
    ties = 0
    for a in range(p):
        for b in range(a + 1, p):
            stat, df = _pearson(data, int(dims[b]), int(dims[a]), b, a, (), dims)
            pv = 1.0 if df < 1 else _chi2.sf(stat, df)
            ties += abs(pv - alpha) < margin
            if pv >= alpha:
                G[a, b] = G[b, a] = False
                sep[(a, b)] = ()
    for l in range(1, max_level + 1):
        nbrs = [list(np.flatnonzero(G[r])) for r in range(p)]
        if max(len(x) for x in nbrs) - 1 < l:
            break
        removed = []
        for a in range(p):
            for b in nbrs[a]:
                if b <= a:
                    continue
                found = None
                for vi, vj in ((a, b), (b, a)):
                    cand = [v for v in nbrs[vi] if v != vj]
                    for S in itertools.combinations(cand, l):
                        stat, df = _pearson(data, int(dims[vi]), int(dims[vj]), vi, vj, S, dims)
                        pv = 1.0 if df < 1 else _chi2.sf(stat, df)
                        ties += abs(pv - alpha) < margin
                        if pv >= alpha and found is None:
                            found = S
                    if found is not None:
                        break
                if found is not None:
                    removed.append((a, b, found))
        for a, b, S in removed:
            G[a, b] = G[b, a] = False
            sep[(a, b)] = tuple(S)
    return G.astype(np.uint16), sep, int(ties)


def gauss_skeleton_pcalg(C: np.ndarray, n: int, alpha: float, max_level: int, margin: float = 1e-9):
    """Textbook restatement of PC-stable with pcalg's conventions: a test of
    (i, j | S) has z = sqrt(n - |S| - 3) atanh|rho| and p = 2 (1 - Phi(z)); the
    edge goes when p >= alpha; pMax[a, b] is the largest p of the tests run on
    the pair; the sepset is recorded for the ordered pair (i, j) whose test
    removed the edge only.  Ordered pairs are tried (a, b) then (b, a), a < b."""
    from scipy.stats import norm

    p = C.shape[0]
    G = np.ones((p, p), bool)
    np.fill_diagonal(G, False)
    sep = {}
    pmax = np.full((p, p), -1.0)
    ties = 0
    q = norm.isf(alpha / 2)

    def test(i, j, S):
        nonlocal ties
        zz = np.sqrt(n - len(S) - 3) * (_z(C, i, j, S) if S else abs(np.arctanh(C[i, j])))
        ties += abs(zz - q) < margin
        a, b = min(i, j), max(i, j)
        pmax[a, b] = max(pmax[a, b], 2 * norm.sf(zz))
        return zz < q

    for a in range(p):
        for b in range(a + 1, p):
            if test(a, b, ()):
                G[a, b] = G[b, a] = False
                sep[(a, b)] = ()
    for l in range(1, max_level + 1):
        nbrs = [list(np.flatnonzero(G[r])) for r in range(p)]
        if max(len(x) for x in nbrs) - 1 < l:
            break
        removed = []
        for a in range(p):
            for b in nbrs[a]:
                if b <= a:
                    continue
                found = None
                for vi, vj in ((a, b), (b, a)):
                    cand = [v for v in nbrs[vi] if v != vj]
                    for S in itertools.combinations(cand, l):
                        if test(vi, vj, S):
                            found = (vi, vj, S)
                            break
                    if found is not None:
                        break
                if found is not None:
                    removed.append(found)
        for vi, vj, S in removed:
            G[vi, vj] = G[vj, vi] = False
            sep[(vi, vj)] = tuple(S)
    return G.astype(np.uint16), sep, pmax, int(ties)
