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
from math import comb

import numpy as np
from scipy.stats import chi2 as _chi2


def _z(C, i, j, S):
    idx = [i, j] + list(S)
    P = np.linalg.inv(C[np.ix_(idx, idx)])
    rho = -P[0, 1] / np.sqrt(P[0, 0] * P[1, 1])
    return abs(np.arctanh(rho))


def gauss_run(C: np.ndarray, thr: np.ndarray, max_level: int, n: int = 4, margin: float = 1e-9) -> dict:
    """Sequential PC-stable (gpucsl conventions) with everything the loop
    observes: skeleton, sepsets {(a, b): S}, zmin (float32, min Fisher z over
    the tests the loop runs on the pair: every test of every level the edge
    survives, then the tests up to the removing one), pmax = 2 (1 - Phi(sqrt(n-3)
    zmin)) with -1 on and below the diagonal, per-level executed counts (tests
    the loop runs: for a < b, (a, b) in rank order up to the first pass, then
    (b, a) the same way if (a, b) kept the edge) and per-level T_l (every
    subset of every ordered pair adjacent at level start)."""
    from scipy.stats import norm

    p = C.shape[0]
    G = np.ones((p, p), bool)
    np.fill_diagonal(G, False)
    sep = {}
    ties = 0
    zmin = np.full((p, p), np.inf)
    executed = np.zeros(max_level + 1, np.int64)
    total = np.zeros(max_level + 1, np.int64)
    for a in range(p):
        for b in range(a + 1, p):
            z = abs(np.arctanh(C[a, b]))
            ties += abs(z - thr[0]) < margin
            zmin[a, b] = min(zmin[a, b], z)
            if z < thr[0]:
                G[a, b] = G[b, a] = False
                sep[(a, b)] = ()
    executed[0] = total[0] = p * (p - 1) // 2
    for l in range(1, max_level + 1):
        nbrs = [list(np.flatnonzero(G[r])) for r in range(p)]
        total[l] = sum(len(x) * comb(len(x) - 1, l) for x in nbrs if len(x) > 0)
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
                        executed[l] += 1
                        ties += abs(z - thr[l]) < margin
                        if not np.isnan(z):
                            zmin[a, b] = min(zmin[a, b], z)
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
    z32 = zmin.astype(np.float32)
    pmax = 2.0 * norm.sf(np.sqrt(n - 3) * z32.astype(np.float64))
    pmax[np.tril_indices(p)] = -1.0
    return dict(skeleton=G.astype(np.uint16), sepsets=sep, zmin=z32, pmax=pmax, executed=executed,
                total=total, ties=int(ties))


def gauss_skeleton(C: np.ndarray, thr: np.ndarray, max_level: int, margin: float = 1e-9):
    r = gauss_run(C, thr, max_level, margin=margin)
    return r["skeleton"], r["sepsets"], r["ties"]


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


def disc_run(data: np.ndarray, dims: np.ndarray, alpha: float, max_level: int, margin: float = 1e-9) -> dict:
    """Sequential discrete PC-stable (gpucsl conventions) on column-major (p, n)
    integer codes.  Every subset of both ordered pairs of an edge adjacent at
    level start is tested (the reference's early exit is a `continue`,
    gpucsl/pc/cuda/discrete_ci.cu:240-243); the sepset is the first passing
    test in pcalg order; pmax[a, b] (a < b) is the max over all those tests of
    the p-value, rounded to float (the reference stores it in a float at every
    level, discrete_ci.cu:90,213,236) and level >= 1 decides on that float
    value; -1 on and below the diagonal.  executed[l] counts the tests run."""
    p = data.shape[0]
    G = np.ones((p, p), bool)
    np.fill_diagonal(G, False)
    sep = {}
    ties = 0
    pmax = np.zeros((p, p), np.float32)
    executed = np.zeros(max_level + 1, np.int64)
    for a in range(p):
        for b in range(a + 1, p):
            stat, df = _pearson(data, int(dims[b]), int(dims[a]), b, a, (), dims)
            pv = 1.0 if df < 1 else _chi2.sf(stat, df)
            ties += abs(pv - alpha) < margin
            pmax[a, b] = max(pmax[a, b], np.float32(pv))
            if pv >= alpha:
                G[a, b] = G[b, a] = False
                sep[(a, b)] = ()
    executed[0] = p * (p - 1) // 2
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
                        pf = np.float32(1.0 if df < 1 else _chi2.sf(stat, df))
                        executed[l] += 1
                        ties += abs(float(pf) - alpha) < margin
                        pmax[a, b] = max(pmax[a, b], pf)
                        if float(pf) >= alpha and found is None:
                            found = S
                if found is not None:
                    removed.append((a, b, found))
        for a, b, S in removed:
            G[a, b] = G[b, a] = False
            sep[(a, b)] = tuple(S)
    pm = pmax.astype(np.float64)
    pm[np.tril_indices(p)] = -1.0
    return dict(skeleton=G.astype(np.uint16), sepsets=sep, pmax=pm, executed=executed, ties=int(ties))


def disc_skeleton(data: np.ndarray, dims: np.ndarray, alpha: float, max_level: int,
                  margin: float = 1e-9):
    """data: column-major (p, n) integer codes."""
    r = disc_run(data, dims, alpha, max_level, margin)
    return r["skeleton"], r["sepsets"], r["ties"]


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
