"""The oracle against closed forms, brute force and invariants (CPU)."""
import numpy as np
import pytest
from scipy import stats

import oracle
from oracle import mini
from gpucsl_amd import synthetic as syn


def test_thresholds_closed_form():
    n, alpha = 2000, 0.01
    thr = oracle.thresholds(n, alpha)
    z = stats.norm.isf(alpha / 2)
    for l in range(48):
        assert thr[l] == pytest.approx(z / np.sqrt(n - l - 4), rel=1e-14)
    assert thr[48] == 0.0 and thr[49] == 0.0
    small = oracle.thresholds(10, alpha)          # min(50, n-3) - 2 = 5 levels filled
    assert np.count_nonzero(small) == 5


def test_gauss_z0_is_atanh():
    for r in np.linspace(-0.999, 0.999, 101):
        assert oracle.gauss_z0(r) == pytest.approx(abs(np.arctanh(r)), rel=1e-12, abs=1e-15)


@pytest.mark.parametrize("level", [1, 2, 3, 4])
def test_gauss_z_is_partial_correlation(level):
    rng = np.random.default_rng(10 + level)
    d = level + 2
    for _ in range(200):
        A = rng.standard_normal((5 * d, d))
        C = np.corrcoef(A, rowvar=False)
        P = np.linalg.inv(C)
        rho = -P[0, 1] / np.sqrt(P[0, 0] * P[1, 1])
        assert oracle.gauss_z(C, 0, 1, list(range(2, d))) == pytest.approx(abs(np.arctanh(rho)), rel=1e-9)


def test_pmax_matches_scipy():
    rng = np.random.default_rng(3)
    z = rng.uniform(0, 0.2, (30, 30)).astype(np.float32)
    n = 5000
    got = oracle.pmax_from_zmin(z, n)
    want = (2 * stats.norm.sf(np.sqrt(n - 3) * z.astype(np.float64))).astype(np.float32)
    iu = np.triu_indices(30, 1)
    assert np.allclose(got[iu], want[iu], rtol=1e-6, atol=1e-30)
    assert (got[np.tril_indices(30)] == -1).all()


def test_pchisq_close_to_scipy():
    # Perlman's approximation: exact series for even df, polynomial normal CDF for odd df
    for df in [1, 2, 3, 4, 7, 10, 31, 100, 501]:
        for x in np.linspace(0.01, 3 * df + 10, 60):
            assert oracle.pchisq(x, df) == pytest.approx(stats.chi2.sf(x, df), abs=2e-6)


def test_corrcoef_matches_numpy():
    X = syn.gaussian_sem(37, 501, 3.0, seed=5).data
    assert np.abs(oracle.corrcoef(X) - np.corrcoef(X, rowvar=False)).max() < 1e-13


def _check_invariants(res, max_level):
    G, sep = res["skeleton"], res["sepsets"]
    p = G.shape[0]
    assert (G == G.T).all() and (np.diag(G) == 0).all()
    assert (sep == np.transpose(sep, (1, 0, 2))).all()
    for a in range(p):
        for b in range(p):
            s = sep[a, b]
            k = int((s >= 0).sum())
            assert (s[:k] >= 0).all() and (s[k:] == -1).all()
            if k:
                assert G[a, b] == 0 and a not in s[:k] and b not in s[:k]
                assert (np.diff(s[:k]) > 0).all()


@pytest.mark.parametrize("seed", range(6))
def test_gauss_skeleton_vs_bruteforce(seed):
    d = syn.gaussian_sem(10, 300, 3.0, seed=seed)
    C = oracle.corrcoef(d.data)
    thr = oracle.thresholds(300, 0.05)
    res = oracle.gauss_skeleton(C, 300, 0.05, 3)
    G, sep, ties = mini.gauss_skeleton(C, thr, 3)
    assert ties == 0
    assert (res["skeleton"] == G).all()
    for (a, b), S in sep.items():
        got = res["sepsets"][a, b]
        assert tuple(got[got >= 0]) == tuple(S), (a, b)
    _check_invariants(res, 3)


def test_gauss_special_cases():
    rng = np.random.default_rng(0)
    n = 4000
    x = rng.standard_normal(n)
    y = 0.9 * x + rng.standard_normal(n) * 0.5          # chain x -> y -> z
    z = 0.9 * y + rng.standard_normal(n) * 0.5
    u = rng.standard_normal(n)                           # collider u -> w <- x
    w = 0.8 * u + 0.8 * x + rng.standard_normal(n) * 0.5
    X = np.stack([x, y, z, u, w], 1)
    res = oracle.gauss_skeleton(oracle.corrcoef(X), n, 0.01, 3)
    G, sep = res["skeleton"], res["sepsets"]
    assert G[0, 1] and G[1, 2] and not G[0, 2]
    assert sep[0, 2, 0] == 1                              # x _||_ z | y
    assert G[3, 4] and G[0, 4] and not G[0, 3]
    assert (sep[0, 3] == -1).all()                        # removed at level 0
    # duplicated column: singular conditioning set -> NaN -> never removed by it
    Xd = np.concatenate([X, X[:, :1]], 1)
    rd = oracle.gauss_skeleton(oracle.corrcoef(Xd), n, 0.01, 2)
    assert rd["skeleton"][0, 5] == 1


def test_gauss_relabel_invariant_skeleton():
    d = syn.gaussian_sem(14, 800, 3.0, seed=11)
    C = oracle.corrcoef(d.data)
    res = oracle.gauss_skeleton(C, 800, 0.01, 3)
    perm = np.random.default_rng(2).permutation(14)
    Cp = C[np.ix_(perm, perm)]
    rp = oracle.gauss_skeleton(np.ascontiguousarray(Cp), 800, 0.01, 3)
    assert (rp["skeleton"] == res["skeleton"][np.ix_(perm, perm)]).all()


def test_gauss_counts_match_definition():
    d = syn.gaussian_sem(25, 1000, 3.0, seed=4)
    res = oracle.gauss_skeleton(oracle.corrcoef(d.data), 1000, 0.01, 3)
    assert res["executed"][0] == res["total"][0] == 25 * 24 // 2
    assert (res["executed"] <= res["total"]).all()


@pytest.mark.parametrize("seed", range(4))
def test_disc_skeleton_vs_bruteforce(seed):
    d = syn.discrete_bn(8, 3000, 2.0, seed=seed)
    data = np.ascontiguousarray(d.data.T)
    dims = (data.max(axis=1) + 1).astype(np.uint8)
    res = oracle.disc_skeleton(data, dims, 0.05, 2)
    G, sep, ties = mini.disc_skeleton(data, dims, 0.05, 2, margin=1e-5)
    assert ties == 0
    assert (res["skeleton"] == G).all()
    for (a, b), S in sep.items():
        got = res["sepsets"][a, b]
        assert tuple(got[got >= 0]) == tuple(S), (a, b)
    _check_invariants(res, 2)
    assert (res["executed"] == res["total"]).all()   # discrete runs the full enumeration


def test_disc_level0_stat_is_pearson():
    rng = np.random.default_rng(7)
    a = rng.integers(0, 4, 500).astype(np.uint8)
    b = ((a + rng.integers(0, 2, 500)) % 3).astype(np.uint8)
    data = np.stack([b, a])
    dims = np.array([3, 4], np.uint8)
    chi2, df, _ = oracle.disc_chi2(data, dims, 1, 0, [])
    table = np.zeros((4, 3))
    np.add.at(table, (a, b), 1)
    want = stats.chi2_contingency(table, correction=False)[0]
    assert chi2 == pytest.approx(want, rel=1e-12)
    assert df == 6


def test_pcalg_mode_matches_textbook_restatement():
    """The oracle's pcalg mode (thresholds sqrt(n-l-3), pMax over every test,
    one-way sepsets) against the brute-force restatement in oracle/mini.py."""
    import oracle
    from oracle import mini
    from gpucsl_amd import synthetic as syn

    for seed in range(4):
        d = syn.gaussian_sem(10, 300, 3.0, seed=70 + seed).data
        C = oracle.corrcoef(d)
        got = oracle.gauss_skeleton(C, 300, 0.05, 3, pcalg_compat=True)
        G, sep, pmax, ties = mini.gauss_skeleton_pcalg(C, 300, 0.05, 3)
        assert ties == 0
        assert (got["skeleton"] == G).all()
        p = 10
        for a in range(p):
            for b in range(p):
                if (a, b) in sep and sep[(a, b)] != ():
                    assert list(got["sepsets"][a, b][: len(sep[(a, b)])]) == list(sep[(a, b)])
                elif not ((b, a) in sep and a > b and sep[(b, a)] == ()) and not ((a, b) in sep):
                    assert (got["sepsets"][a, b] == -1).all(), (a, b)
        iu = np.triu_indices(p, 1)
        assert np.abs(got["pmax"][iu] - pmax[iu]).max() < 2e-6


def test_pcalg_thresholds_closed_form():
    import oracle
    from scipy.stats import norm

    thr = oracle.pcalg_thresholds(100, 0.01, 10)
    for l in range(10):
        assert thr[l] == pytest.approx(norm.isf(0.005) / np.sqrt(100 - l - 3), rel=1e-14)


@pytest.mark.parametrize("seed", range(4))
def test_gauss_pmax_and_counts_vs_bruteforce(seed):
    """zmin/pmax (min over exactly the tests a sequential pcalg-order loop runs)
    and the per-level executed and T_l counts of the C oracle against the
    brute-force loop, which counts every test it runs."""
    n = 400
    d = syn.gaussian_sem(10, n, 4.0, seed=30 + seed)
    C = oracle.corrcoef(d.data)
    thr = oracle.thresholds(n, 0.05)
    res = oracle.gauss_skeleton(C, n, 0.05, 4)
    ref = mini.gauss_run(C, thr, 4, n=n)
    assert ref["ties"] == 0
    assert (res["skeleton"] == ref["skeleton"]).all()
    L = len(ref["executed"])
    assert (res["executed"][:L] == ref["executed"]).all(), (res["executed"], ref["executed"])
    assert (res["total"][:L] == ref["total"]).all(), (res["total"], ref["total"])
    assert ref["executed"][1] > 0
    iu = np.triu_indices(10, 1)
    # two FP64 formulas of the same statistic, rounded to float32 zmin
    assert np.allclose(res["pmax"][iu], ref["pmax"][iu], rtol=1e-5, atol=1e-9)
    assert (res["pmax"][np.tril_indices(10)] == -1).all()


@pytest.mark.parametrize("seed", range(3))
def test_disc_pmax_and_counts_vs_bruteforce(seed):
    """Discrete pmax is the max, over every subset of both ordered pairs of an
    edge adjacent at level start, of the float-rounded p-value; executed is
    that full enumeration.  Checked against the brute-force loop (scipy's
    chi2.sf; the oracle uses Perlman's pchisq, within 2e-6 of it)."""
    d = syn.discrete_bn(8, 3000, 2.5, seed=20 + seed)
    data = np.ascontiguousarray(d.data.T)
    dims = (data.max(axis=1) + 1).astype(np.uint8)
    res = oracle.disc_skeleton(data, dims, 0.05, 3)
    ref = mini.disc_run(data, dims, 0.05, 3, margin=1e-5)
    assert ref["ties"] == 0
    assert (res["skeleton"] == ref["skeleton"]).all()
    L = len(ref["executed"])
    assert (res["executed"][:L] == ref["executed"]).all(), (res["executed"], ref["executed"])
    assert ref["executed"][1] > 0
    iu = np.triu_indices(8, 1)
    assert np.abs(res["pmax"][iu] - ref["pmax"][iu]).max() < 4e-6
    assert (res["pmax"][np.tril_indices(8)] == -1).all()
