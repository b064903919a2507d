"""Discrete (Pearson chi^2) skeleton on the GPU against the FP64 oracle and the
reference's known-answer vectors (GPU box only)."""
import itertools
import sys

import numpy as np
import pytest

import oracle
from gpucsl_amd import _native as nat
from gpucsl_amd import synthetic as syn
from gpucsl_amd.cli.cli_util import GpucslError

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _engine(codes, max_level=1, alpha=0.05):
    from gpucsl_amd.discrete import DiscreteEngine

    codes = np.asarray(codes, np.uint8)
    n, p = codes.shape
    dims = (codes.max(axis=0).astype(np.int64) + 1).astype(np.uint8)
    eng = DiscreteEngine(p, n, max_level, alpha, int(dims.max()))
    eng.set_codes(codes, dims)
    return eng, dims


def test_contingency_level0_golden(golden):
    c = golden["contingency_level0"]
    codes = np.stack([c["a"], c["b"]], axis=1)
    eng, _ = _engine(codes)
    table, chi2, p, df = eng.ci_test(0, 1, [])
    assert table.reshape(7, 2).tolist() == c["expected"]
    assert df == 6


def test_contingency_level2_golden(golden):
    c = golden["contingency_level2"]
    eng, _ = _engine(np.array(c["data"]))
    table, _, _, df = eng.ci_test(c["vi"], c["vj"], c["sepset"])
    assert table.tolist() == c["expected"]
    assert df == 4 * 4 * 25


def test_chi2_golden_and_pchisq(golden):
    """The reference's table of tests/test_cuda_discrete_helpers.py:128-172,
    rebuilt as observations (stratum s, v_i = a, v_j = b)."""
    c = golden["marginals_chi2"]
    N = np.array(c["N"]).reshape(5, 5, 5)
    rows = [(s, a, b) for s, a, b in itertools.product(range(5), repeat=3) for _ in range(N[s, a, b])]
    codes = np.array(rows, np.uint8)
    eng, _ = _engine(codes)
    table, chi2, p, df = eng.ci_test(1, 2, [0])
    assert table.tolist() == N.flatten().tolist()
    assert abs(chi2 - c["chi2"]) <= c["tol"]
    assert df == 80
    assert p == pytest.approx(oracle.pchisq(chi2, df), rel=1e-14, abs=1e-300)


@pytest.mark.parametrize("l", [0, 1, 2, 3])
def test_single_tests_match_oracle(l):
    d = syn.discrete_bn(12, 3001, 2.0, seed=40 + l)
    codes = d.data
    eng, dims = _engine(codes)
    colmajor = np.ascontiguousarray(codes.T)
    rng = np.random.default_rng(l)
    for _ in range(20):
        vs = rng.choice(12, size=l + 2, replace=False)
        vi, vj, S = int(vs[0]), int(vs[1]), [int(v) for v in sorted(vs[2:])]
        table, chi2, p, df = eng.ci_test(vi, vj, S)
        ochi2, odf, work = oracle.disc_chi2(colmajor, dims, vi, vj, S)
        assert table.tolist() == work[: table.size].tolist()
        assert df == odf
        assert chi2 == ochi2 or abs(chi2 - ochi2) <= 1e-12 * abs(ochi2)
        assert p == pytest.approx(oracle.pchisq(ochi2, odf), rel=1e-13, abs=1e-300)


def test_large_df_pvalues_match_oracle():
    """Level-3 tests with 5 categories: df = 4 * 4 * 125 = 2000, so the
    p-value series runs a thousand iterations (evaluated by a whole wave)."""
    d = syn.discrete_bn(10, 4000, 2.0, seed=91, card_low=5, card_high=5)
    eng, dims = _engine(d.data)
    colmajor = np.ascontiguousarray(d.data.T)
    rng = np.random.default_rng(91)
    dfs = []
    for _ in range(12):
        vs = rng.choice(10, size=5, replace=False)
        vi, vj, S = int(vs[0]), int(vs[1]), [int(v) for v in sorted(vs[2:])]
        table, chi2, p, df = eng.ci_test(vi, vj, S)
        ochi2, odf, work = oracle.disc_chi2(colmajor, dims, vi, vj, S)
        assert table.tolist() == work[: table.size].tolist() and df == odf
        assert chi2 == ochi2 or abs(chi2 - ochi2) <= 1e-12 * abs(ochi2)
        assert p == pytest.approx(oracle.pchisq(ochi2, odf), rel=1e-12, abs=1e-300)
        dfs.append(df)
    say(f"[large df] df range {min(dfs)}..{max(dfs)}")
    assert max(dfs) >= 2000
    # 6 categories: 7776 cells + marginals exceed the LDS of the single-test entry point
    d6 = syn.discrete_bn(6, 2000, 2.0, seed=92, card_low=6, card_high=6)
    eng6, _ = _engine(d6.data)
    with pytest.raises(GpucslError):
        eng6.ci_test(0, 1, [2, 3, 4])


def _run(codes, alpha, L):
    eng, dims = _engine(codes, L, alpha)
    eng.run()
    res = eng.results()
    ref = oracle.disc_skeleton(np.ascontiguousarray(np.asarray(codes, np.uint8).T), dims, alpha, L)
    return res, ref


def _compare(res, ref, label):
    same_skel = (res.skeleton == ref["skeleton"]).all()
    same_sep = (res.separation_sets == ref["sepsets"]).all()
    iu = np.triu_indices(res.skeleton.shape[0], 1)
    dp = np.abs(res.pmax[iu].astype(np.float64) - ref["pmax"][iu])
    exact = float((res.pmax[iu] == ref["pmax"][iu]).mean()) if iu[0].size else 1.0
    say(f"[{label}] skeleton_equal={same_skel} sepsets_equal={same_sep} pmax_bitexact={exact:.5f} "
        f"max|dpmax|={dp.max() if dp.size else 0:.3g} edges={int(res.skeleton.sum()) // 2}")
    for l, st in enumerate(res.stats):
        say(f"   level {l}: {st} oracle_exec={int(ref['executed'][l])}")
    assert same_skel and same_sep
    assert dp.max() <= 1e-6
    assert (res.pmax[np.tril_indices(res.skeleton.shape[0])] == -1).all()
    for l in range(len(res.stats)):
        assert res.stats[l]["executed"] == ref["executed"][l]


@pytest.mark.parametrize("seed", range(4))
def test_small_parity(seed):
    d = syn.discrete_bn(30, 4000, 2.0, seed=seed)
    res, ref = _run(d.data, 0.05, 3)
    _compare(res, ref, f"disc seed {seed}")


def test_more_variables_than_samples():
    """p = 80 discrete columns from n = 60 rows (sparse contingency tables,
    many empty strata) against the oracle, every level's test count included."""
    d = syn.discrete_bn(80, 60, 2.0, seed=61)
    res, ref = _run(d.data, 0.05, 3)
    _compare(res, ref, "disc p > n")


def test_c3_parity():
    """Config C3: p=200, n=50000, cardinalities 2..5, degree 2, alpha=0.01, L=3 (seed 200)."""
    d = syn.discrete_bn(200, 50000, 2.0, seed=200)
    res, ref = _run(d.data, 0.01, 3)
    _compare(res, ref, "C3")


def test_high_levels_parity():
    """Levels 4-8: a binary network (p=16, n=50 000, degree 5) whose skeleton
    still has tests at level 8."""
    d = syn.discrete_bn(16, 50000, 5.0, seed=2, card_low=2, card_high=2)
    res, ref = _run(d.data, 0.05, 8)
    assert ref["executed"][8] > 0
    _compare(res, ref, "discrete, max level 8")


def test_levels_above_8_parity():
    """Levels 9-10 (the kernel that takes the level at run time): a binary
    variable, ten noisy copies of it and a threshold of their sum, so the edge
    between the first and the last goes at level 10; skeleton, sepsets,
    per-level test counts and pmax against the oracle."""
    rng = np.random.default_rng(5)
    n, k = 50000, 10
    i = rng.integers(0, 2, n)
    m = (i[:, None] ^ (rng.random((n, k)) < 0.2)).astype(np.int64)
    j = ((m.sum(1) + rng.integers(0, 2, n)) > 5).astype(np.int64)
    codes = np.column_stack([i, m, j]).astype(np.uint8)
    res, ref = _run(codes, 0.01, 10)
    assert ref["executed"][9] > 0 and ref["executed"][10] > 0
    assert ref["skeleton"][0, 11] == 0
    _compare(res, ref, "discrete, max level 10")


@pytest.mark.parametrize("p,card,n,label", [
    (14, (14, 16), 2500, "one-hot Gram matrix, 16 categories"),
    (12, (17, 20), 2500, "pair-by-pair tables, more than 16 categories"),
    (300, (14, 14), 1500, "pair-by-pair tables, p * max_dim > 4096"),
])
def test_level0_table_paths(p, card, n, label):
    """Level 0 takes its tables from the int8-MFMA Gram matrix of the one-hot
    codes or counts them pair by pair, by category count and width; both
    match the oracle."""
    d = syn.discrete_bn(p, n, 2.0, seed=70 + p, card_low=card[0], card_high=card[1])
    res, ref = _run(d.data, 0.05, 1)
    _compare(res, ref, label)


def test_level0_runs_again_on_the_same_workspace():
    """The level-0 table matrix is left zero by each run, so level 0 can run
    again after gpucsl_disc_init without double counting."""
    d = syn.discrete_bn(20, 3000, 2.0, seed=12)
    eng, dims = _engine(d.data, 2, 0.05)
    eng.start()
    torch.cuda.synchronize()
    G1, pm1 = eng.G.cpu().numpy().copy(), eng.pmax.cpu().numpy().copy()
    s = nat.stream_handle()
    nat.call("gpucsl_disc_level0", eng.data.data_ptr(), eng.n, eng.stride, eng.dims.data_ptr(), eng.p,
             eng.max_level, eng.max_dim, eng.alpha, eng.G.data_ptr(), eng.pmax.data_ptr(),
             eng.stats.data_ptr(), eng.ws.data_ptr(), eng.ws_bytes, s)
    torch.cuda.synchronize()
    assert (eng.G.cpu().numpy() == G1).all()
    assert np.array_equal(eng.pmax.cpu().numpy().view(np.uint32), pm1.view(np.uint32))


def test_constant_column_and_unaligned_n():
    d = syn.discrete_bn(10, 1237, 2.0, seed=5)
    codes = np.concatenate([d.data, np.zeros((1237, 1), np.uint8)], axis=1)
    res, ref = _run(codes, 0.05, 2)
    _compare(res, ref, "constant column")
    assert res.skeleton[10].sum() == 0          # df = 0 -> p = 1 -> removed at level 0


def test_dropin_api_and_encoding():
    from gpucsl_amd.pc.discover_skeleton_discrete import discover_skeleton_gpu_discrete
    from gpucsl_amd.pc.helpers import init_pc_graph

    d = syn.discrete_bn(15, 3000, 2.0, seed=8)
    raw = d.data.astype(np.int64) * 7 + 3          # non-contiguous category labels
    res, runtime = discover_skeleton_gpu_discrete(init_pc_graph(raw), raw, 0.05, 2, 15, 3000)
    assert runtime > 0 and res.discover_skeleton_kernel_time > 0
    dims = (d.data.max(axis=0) + 1).astype(np.uint8)
    ref = oracle.disc_skeleton(np.ascontiguousarray(d.data.T), dims, 0.05, 2)
    assert (res.skeleton == ref["skeleton"]).all()
    assert (res.seperation_sets == ref["sepsets"]).all()
    with pytest.raises(RuntimeError):
        discover_skeleton_gpu_discrete(init_pc_graph(raw), raw, 0.05, 2, 15, 3000, memory_restriction=10)


def test_skeleton_twice_in_a_row():
    """The whole level-1 skeleton twice on one engine (its buffers and workspace
    left from the first run) and twice through the drop-in API, as the
    reference's tests/test_skeleton_discrete.py:96-108 runs it: every run equals
    the oracle, including the per-level test counts and pmax."""
    from gpucsl_amd.pc.discover_skeleton_discrete import discover_skeleton_gpu_discrete
    from gpucsl_amd.pc.helpers import init_pc_graph

    d = syn.discrete_bn(25, 4000, 2.0, seed=96)
    dims = (d.data.max(axis=0) + 1).astype(np.uint8)
    ref = oracle.disc_skeleton(np.ascontiguousarray(d.data.T), dims, 0.05, 1)
    eng, _ = _engine(d.data, 1, 0.05)
    for k in range(2):
        eng.run()
        _compare(eng.results(), ref, f"same engine, run {k + 1}")
    raw = d.data.astype(np.int64)
    for k in range(2):
        res, _ = discover_skeleton_gpu_discrete(init_pc_graph(raw), raw, 0.05, 1, 25, 4000)
        say(f"[drop-in run {k + 1}] edges={int(res.skeleton.sum()) // 2}")
        assert (res.skeleton == ref["skeleton"]).all()
        assert (res.seperation_sets == ref["sepsets"]).all()


def test_memory_restriction_runs_fewer_tests_at_once_same_result():
    """Tables of max_dim=8 at level 3 exceed the LDS budget; a workspace with
    room for three global table slices runs three tests at a time and must
    give the same result as the full workspace."""
    from gpucsl_amd import _native as nat
    from gpucsl_amd.pc.discover_skeleton_discrete import discover_skeleton_gpu_discrete
    from gpucsl_amd.pc.helpers import init_pc_graph

    d = syn.discrete_bn(12, 3000, 2.0, seed=31, card_low=6, card_high=8)
    raw = d.data.astype(np.int64)
    lib = nat.load()
    full = int(lib.gpucsl_disc_workspace_bytes(12, 3, 8))
    least = int(lib.gpucsl_disc_min_workspace_bytes(12, 3, 8))
    slice_bytes = (full - least) // max(1, 1023)   # grid of 1024 blocks -> 1023 extra slices
    assert least < full
    a, _ = discover_skeleton_gpu_discrete(init_pc_graph(raw), raw, 0.05, 3, 12, 3000)
    b, _ = discover_skeleton_gpu_discrete(init_pc_graph(raw), raw, 0.05, 3, 12, 3000,
                                          memory_restriction=least + 2 * slice_bytes)
    say(f"workspace full {full} min {least}")
    assert (a.skeleton == b.skeleton).all() and (a.seperation_sets == b.seperation_sets).all()
    assert np.array_equal(a.pmax.view(np.uint32), b.pmax.view(np.uint32))


@pytest.mark.parametrize("dtype", [np.int64, np.int32, np.int16, np.uint8, np.bool_, np.float64, np.float32])
def test_device_ordinal_encoding_matches_host(dtype):
    """gpucsl_disc_encode gives the host encoder's codes and category counts
    (sorted distinct values -> 0..k-1) for each numeric dtype."""
    from gpucsl_amd.discrete import ordinal_encode, ordinal_encode_device

    rng = np.random.default_rng(7)
    n, p = 1237, 9
    if dtype == np.bool_:
        raw = rng.integers(0, 2, size=(n, p)).astype(np.bool_)
    elif np.issubdtype(dtype, np.floating):
        raw = (rng.integers(-3, 4, size=(n, p)) * 0.37 + 1e-3).astype(dtype)
    elif dtype == np.uint8:
        raw = (rng.integers(0, 6, size=(n, p)) * 40 + 3).astype(dtype)
    else:
        raw = (rng.integers(-4, 5, size=(n, p)) * 1000 - 17).astype(dtype)
    raw[:, 3] = raw[0, 3]                      # a constant column
    want_codes, want_dims = ordinal_encode(raw)
    codes_t, dims_t, dims = ordinal_encode_device(raw)
    codes = codes_t.cpu().numpy()
    assert (dims == want_dims).all()
    assert (codes[:, :n].T == want_codes).all()
    assert (codes[:, n:] == 0).all()


def test_device_ordinal_encoding_limits():
    from gpucsl_amd.discrete import ordinal_encode_device

    wide = np.tile(np.arange(300, dtype=np.int64)[:, None], (1, 3))
    with pytest.raises(GpucslError):
        ordinal_encode_device(wide)
    with_nan = np.ones((50, 2))
    with_nan[4, 1] = np.nan
    assert ordinal_encode_device(with_nan) is None        # left to the host encoder
    assert ordinal_encode_device(np.array([["a", "b"], ["b", "a"]])) is None


def test_device_ordinal_encoding_column_major_input():
    """gpucsl_disc_encode on column-major values with a padded leading
    dimension gives the same codes as on the row-major layout."""
    from gpucsl_amd.discrete import ordinal_encode

    rng = np.random.default_rng(11)
    n, p, ld = 999, 5, 1024
    raw = rng.integers(-2, 3, size=(n, p)).astype(np.int64) * 7
    cm = np.zeros((p, ld), np.int64)
    cm[:, :n] = raw.T
    vals = torch.from_numpy(cm).cuda()
    stride = 1024
    codes = torch.empty((p, stride), dtype=torch.uint8, device="cuda")
    dims = torch.empty((p,), dtype=torch.uint8, device="cuda")
    status = torch.empty((p,), dtype=torch.int32, device="cuda")
    nat.call("gpucsl_disc_encode", vals.data_ptr(), 1, n, p, ld, 0, codes.data_ptr(), stride, dims.data_ptr(),
             status.data_ptr(), nat.stream_handle())
    want_codes, want_dims = ordinal_encode(raw)
    assert (status.cpu().numpy() == 0).all()
    assert (dims.cpu().numpy() == want_dims).all()
    assert (codes.cpu().numpy()[:, :n].T == want_codes).all()
