"""The oracle's restatement against the reference helpers themselves, compiled
for the host from the reference checkout into oracle/_ref (CPU).  Bit-for-bit
where the reference code is deterministic."""
import ctypes as ct

import numpy as np
import pytest

import oracle

R = oracle.ref()
pytestmark = pytest.mark.skipif(R is None, reason="oracle/_ref not built (no reference checkout)")


def test_binomial_matches_ref():
    for n in range(0, 40):
        for k in range(0, 12):
            if n >= k and oracle.binomial(n, k) * max(k, 1) >= 2**31:
                continue  # the reference's int overflows there
            assert oracle.binomial(n, k) == R.ref_binomial(n, k), (n, k)


def test_poz_bitwise():
    for z in np.concatenate([np.linspace(-7, 7, 2001), [0.0, 1e-300, -1e-300, 6.0, -6.0, 2.0]]):
        assert oracle.poz(float(z)) == R.ref_poz(float(z))


def test_pchisq_bitwise():
    rng = np.random.default_rng(1)
    for df in list(range(1, 60)) + [99, 100, 250, 999, 1000, 2000, 3125]:
        for x in np.concatenate([rng.uniform(0, 3 * df, 40), [0.0, 1e-9, 39.9, 40.0, 40.1]]):
            a, b = oracle.pchisq(float(x), df), R.ref_pchisq(float(x), df)
            assert a == b or (a != a and b != b), (x, df, a, b)


@pytest.mark.parametrize("level", [1, 2, 3, 4, 5])
def test_unrank_matches_ref(level):
    out = np.zeros(8, np.int32)
    for N in range(level, 14):
        total = oracle.binomial(N, level)
        for r in range(total):
            R.ref_sepset_indices(level, N, r, out.ctypes.data_as(ct.c_void_p))
            assert oracle.unrank(N, level, r).tolist() == out[:level].tolist()


def _random_corr(d, rng, cond=1.0):
    A = rng.standard_normal((d * 4, d)) @ np.diag(rng.uniform(0.2, 1.0, d) ** cond)
    A = A + rng.standard_normal((d * 4, 1)) * rng.uniform(0, 2)
    return np.corrcoef(A, rowvar=False)


@pytest.mark.parametrize("level", [1, 2, 3, 4, 5])
def test_gauss_z_bitwise(level):
    rng = np.random.default_rng(level)
    d = level + 2
    for t in range(300):
        C = np.ascontiguousarray(_random_corr(d, rng))
        want = R.ref_gauss_z_packed(level, C.ctypes.data_as(ct.c_void_p))
        got = oracle.gauss_z(C, 0, 1, list(range(2, d)))
        assert got == want or (got != got and want != want), (t, got, want)


def test_gauss_z_singular_matches_ref():
    # duplicated conditioning variable: M2 singular -> NaN in both
    C = np.eye(4)
    C[0, 1] = C[1, 0] = 0.3
    C[2, 3] = C[3, 2] = 1.0
    want = R.ref_gauss_z_packed(2, C.ctypes.data_as(ct.c_void_p))
    got = oracle.gauss_z(C, 0, 1, [2, 3])
    assert (got != got and want != want) or got == want


def test_matrix_helpers_vs_ref(golden):
    f = {"inverse3": R.ref_inverse3, "cholesky3": R.ref_cholesky3, "m2_inv3": R.ref_m2_inv3}
    for key, fn in f.items():
        A = np.ascontiguousarray(np.array(golden[key]["A"], float))
        out = np.zeros(9)
        fn(A.ctypes.data_as(ct.c_void_p), out.ctypes.data_as(ct.c_void_p))
        assert np.allclose(out.reshape(3, 3), golden[key]["expected"]), key
