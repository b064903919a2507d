"""The oracle against the reference's own known-answer vectors (CPU)."""
import numpy as np
import pytest


def test_binomial(golden, oracle_mod):
    for n, k, want in golden["binomial"]:
        assert oracle_mod.binomial(n, k) == want


def test_pchisq_float_eq(golden, oracle_mod):
    # tests/cuda/helper_test.cc uses EXPECT_FLOAT_EQ: equal within 4 float ulps
    for x, df, want in golden["pchisq"]:
        got = np.float32(oracle_mod.pchisq(x, df))
        w = np.float32(want)
        ulps = abs(int(got.view(np.int32)) - int(w.view(np.int32)))
        assert ulps <= 4, (x, df, got, w)


def test_unrank_5c3(golden, oracle_mod):
    for rank, want in enumerate(golden["unrank_5c3"]):
        assert oracle_mod.unrank(5, 3, rank).tolist() == want


@pytest.mark.parametrize("case", ["compact_random", "compact_full"])
def test_compact(golden, oracle_mod, case):
    G = np.array(golden[case]["G"], np.uint16)
    assert oracle_mod.compact(G).tolist() == golden[case]["Gc"]


def test_contingency_level0(golden, oracle_mod):
    c = golden["contingency_level0"]
    a, b = np.array(c["a"], np.uint8), np.array(c["b"], np.uint8)
    data = np.stack([b, a])            # column-major (p, n): variable 0 = b, 1 = a
    dims = np.array([2, 7], np.uint8)
    _, df, work = oracle_mod.disc_chi2(data, dims, 1, 0, [])
    table = work[:14].reshape(7, 2)
    assert table.tolist() == c["expected"]
    assert df == 6


def test_contingency_level2(golden, oracle_mod):
    c = golden["contingency_level2"]
    data = np.ascontiguousarray(np.array(c["data"], np.uint8).T)
    dims = np.array(c["dims"], np.uint8)
    _, df, work = oracle_mod.disc_chi2(data, dims, c["vi"], c["vj"], c["sepset"])
    assert work[:625].tolist() == c["expected"]
    assert df == 4 * 4 * 25


def test_marginals_and_chi2(golden, oracle_mod):
    c = golden["marginals_chi2"]
    got = oracle_mod.chi2_table(c["N"], c["Nis"], c["Njs"], c["Ns"], c["dim_s"], c["di"], c["dj"])
    assert abs(got - c["chi2"]) <= c["tol"]
    got_seq = oracle_mod.chi2_table(c["N"], c["Nis"], c["Njs"], c["Ns"], c["dim_s"], c["di"], c["dj"],
                                    sequential=True)
    assert abs(got_seq - c["chi2"]) <= c["tol"]
