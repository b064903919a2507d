"""Gaussian skeleton on the GPU against the FP64 oracle (GPU box only)."""
import sys

import numpy as np
import pytest

import oracle
from gpucsl_amd import _native as nat
from gpucsl_amd import synthetic as syn

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    nat.load()
    yield
    torch.cuda.synchronize()


def _engine_run(X, n, alpha, L, C=None, margin=None, group_size=None):
    from gpucsl_amd.gaussian import DEFAULT_GROUP_SIZE, DEFAULT_MARGIN, GaussianEngine
    from gpucsl_amd.pc.helpers import get_gaussian_thresholds

    p = X.shape[1] if X is not None else C.shape[0]
    thr = get_gaussian_thresholds(np.zeros((n, 1)), alpha)
    eng = GaussianEngine(p, n, L, thr, margin=margin or DEFAULT_MARGIN,
                         group_size=group_size or DEFAULT_GROUP_SIZE)
    if C is None:
        eng.set_data(X)
    else:
        eng.set_correlation(C)
    eng.run()
    res = eng.results()
    torch.cuda.synchronize()
    return eng, res


def _compare(res, ref, label, pmax_atol=2e-5):
    same_skel = (res.skeleton == ref["skeleton"]).all()
    same_sep = (res.separation_sets == ref["sepsets"]).all()
    iu = np.triu_indices(res.skeleton.shape[0], 1)
    dp = float(np.abs(res.pmax[iu].astype(np.float64) - ref["pmax"][iu]).max()) if iu[0].size else 0.0
    say(f"[{label}] skeleton_equal={same_skel} sepsets_equal={same_sep} max|dpmax|={dp:.3g} "
        f"edges={int(res.skeleton.sum()) // 2}")
    for l, st in enumerate(res.stats):
        say(f"   level {l}: {st} oracle_exec={int(ref['executed'][l])} oracle_total={int(ref['total'][l])}")
    assert same_skel
    assert same_sep
    assert dp <= pmax_atol
    assert ((res.pmax[np.tril_indices(res.skeleton.shape[0])]) == -1).all()


def test_unrank_matches_oracle():
    for l in list(range(1, 6)) + [9, 10, 12]:
        for N in sorted({l, l + 1, 7, 13, 40, 150}):
            total = oracle.binomial(N, l)
            if total == 0:
                continue
            if total <= 4000:
                ranks = np.arange(total, dtype=np.uint64)
            else:
                rng = np.random.default_rng(N * 10 + l)
                ranks = np.unique(np.concatenate([rng.integers(0, total, 3000), [0, total - 1]])).astype(np.uint64)
            d_r = torch.from_numpy(ranks.view(np.int64)).cuda()
            d_o = torch.empty((ranks.size * l,), dtype=torch.int32, device="cuda")
            nat.call("gpucsl_unrank", N, l, d_r.data_ptr(), ranks.size, d_o.data_ptr(), nat.stream_handle())
            got = d_o.cpu().numpy().reshape(-1, l)
            for k, r in enumerate(ranks):
                assert got[k].tolist() == oracle.unrank(N, l, int(r)).tolist(), (N, l, int(r))


@pytest.mark.parametrize("case", ["compact_random", "compact_full"])
def test_compact_golden(golden, case):
    G = torch.tensor(golden[case]["G"], dtype=torch.uint8, device="cuda")
    p = G.shape[0]
    ws_bytes = nat.load().gpucsl_gauss_workspace_bytes(p, 1)
    ws = torch.empty((ws_bytes,), dtype=torch.uint8, device="cuda")
    Gc = torch.empty((p, p), dtype=torch.int32, device="cuda")
    nat.call("gpucsl_compact", G.data_ptr(), p, Gc.data_ptr(), ws.data_ptr(), ws_bytes, nat.stream_handle())
    assert Gc.cpu().numpy().tolist() == golden[case]["Gc"]


def test_compact_random_large():
    rng = np.random.default_rng(0)
    p = 777
    G = (rng.random((p, p)) < 0.3).astype(np.uint8)
    G = np.maximum(G, G.T)
    d_G = torch.from_numpy(G).cuda()
    ws_bytes = nat.load().gpucsl_gauss_workspace_bytes(p, 1)
    ws = torch.empty((ws_bytes,), dtype=torch.uint8, device="cuda")
    Gc = torch.empty((p, p), dtype=torch.int32, device="cuda")
    nat.call("gpucsl_compact", d_G.data_ptr(), p, Gc.data_ptr(), ws.data_ptr(), ws_bytes, nat.stream_handle())
    assert (Gc.cpu().numpy() == oracle.compact(G.astype(np.uint16))).all()


@pytest.mark.parametrize("p,n,dtype", [(50, 2000, np.float64), (64, 1000, np.float32),
                                       (65, 333, np.float32), (130, 2049, np.float64),
                                       (256, 3001, np.float32), (257, 3001, np.float32),
                                       (300, 2500, np.float64), (1000, 10000, np.float32)])
def test_corr_matches_numpy(p, n, dtype):
    X = syn.gaussian_sem(p, n, 3.0, seed=p + n).data.astype(dtype)
    from gpucsl_amd.gaussian import GaussianEngine

    eng = GaussianEngine(p, n, 1, [0.1, 0.1])
    eng.set_data(X)
    C64 = eng.C64.cpu().numpy()
    C32 = eng.C32.cpu().numpy()
    want = np.corrcoef(X.astype(np.float64), rowvar=False)
    err = float(np.abs(C64 - want).max())
    say(f"corr p={p} n={n} {np.dtype(dtype).name}: max|C64 - numpy| = {err:.3g}")
    assert err < 1e-12
    assert (C32 == C64.astype(np.float32)).all()
    assert (C64 == C64.T).all()


def _corr_engine(X):
    from gpucsl_amd.gaussian import GaussianEngine

    eng = GaussianEngine(X.shape[1], X.shape[0], 1, [0.1, 0.1])
    eng.set_data(X)
    return eng.C64.cpu().numpy(), eng.C32.cpu().numpy()


def test_corr_float32_nonfinite_and_constant_columns():
    """A column holding NaN or Inf, or a constant column, gets NaN correlations
    in its row and column, as with numpy's FP64 corrcoef; the rest match it."""
    n, p = 3000, 40
    X = syn.gaussian_sem(p, n, 3.0, seed=5).data.astype(np.float32)
    X[17, 3] = np.nan
    X[:, 5] = 2.5
    X[400, 7] = np.inf
    C64, C32 = _corr_engine(X)
    with np.errstate(all="ignore"):
        want = np.corrcoef(X.astype(np.float64), rowvar=False)
    bad = np.zeros(p, bool)
    bad[[3, 5, 7]] = True
    assert np.isnan(C64[bad]).all() and np.isnan(C64[:, bad]).all()
    good = np.ix_(~bad, ~bad)
    assert not np.isnan(C64[good]).any()
    err = float(np.abs(C64[good] - want[good]).max())
    say(f"corr with NaN/Inf/constant columns: max|C64 - numpy| over the rest = {err:.3g}")
    assert err < 1e-12
    assert (C32[good] == C64[good].astype(np.float32)).all()


def test_corr_float32_offsets_and_scales():
    """Columns with a large offset and a small spread, and columns scaled by
    1e-30 and 1e30: each column is scaled by its own power of two before the
    integer Gram matrix, so the correlation keeps FP64 accuracy."""
    n, p = 5000, 24
    X = syn.gaussian_sem(p, n, 3.0, seed=6).data
    X[:, :4] = 1e4 + 1e-2 * X[:, :4]
    X[:, 4:8] *= 1e-30
    X[:, 8:12] *= 1e30
    X = X.astype(np.float32)
    C64, _ = _corr_engine(X)
    want = np.corrcoef(X.astype(np.float64), rowvar=False)
    err = float(np.abs(C64 - want).max())
    say(f"corr offsets/scales: max|C64 - numpy| = {err:.3g}")
    assert err < 1e-12


def test_corr_float32_many_rows_and_split_limit():
    """n = 70 000 rows, p = 300 (float32 data above 256 columns takes the
    integer Gram path): it needs at least two row splits (at most 65 536 rows
    each keeps the int32 sums exact); a job that asks for one is refused."""
    from gpucsl_amd.batch import GaussJobState, GaussianBatch
    from gpucsl_amd.cli.cli_util import GpucslError

    n, p = 70000, 300
    assert nat.load().gpucsl_corr_k_splits(n, p, 0) >= 2
    X = syn.gaussian_sem(p, n, 3.0, seed=8).data.astype(np.float32)
    C64, _ = _corr_engine(X)
    err = float(np.abs(C64 - np.corrcoef(X.astype(np.float64), rowvar=False)).max())
    say(f"corr n={n}: max|C64 - numpy| = {err:.3g}")
    assert err < 1e-12
    job = GaussJobState(n, p, 1, alpha=0.01, k_splits=1)
    job.set_data(torch.from_numpy(X).cuda())
    b = GaussianBatch([job])
    with pytest.raises(GpucslError, match="rows per split"):
        nat.call("gpucsl_gauss_batch_corr", *b._args(), nat.stream_handle())


@pytest.mark.parametrize("seed", range(10))
def test_c1_parity(seed):
    """Config C1: p=50, n=2000 FP64 samples, degree 3, alpha=0.01, L=3."""
    d = syn.gaussian_sem(50, 2000, 3.0, seed=seed)
    _, res = _engine_run(d.data, 2000, 0.01, 3)
    ref = oracle.gauss_skeleton(oracle.corrcoef(d.data), 2000, 0.01, 3)
    _compare(res, ref, f"C1 seed {seed}")
    for l in range(4):
        assert res.stats[l]["useful"] == ref["executed"][l]
        assert res.stats[l]["total"] == ref["total"][l]


@pytest.mark.parametrize("group_size", [8, 32, 64])
def test_group_sizes_agree(group_size):
    d = syn.gaussian_sem(80, 3000, 4.0, seed=7)
    _, res = _engine_run(d.data, 3000, 0.01, 4, group_size=group_size)
    ref = oracle.gauss_skeleton(oracle.corrcoef(d.data), 3000, 0.01, 4)
    _compare(res, ref, f"group {group_size}")


@pytest.mark.parametrize("seed", [100, 101, 102])
def test_c2_parity(seed):
    """Config C2: p=1000, n=10000 FP32, degree 4, alpha=0.01, L=4 (seed 100 is
    the benchmark's dataset)."""
    d = syn.gaussian_sem(1000, 10000, 4.0, seed=seed)
    X = d.data.astype(np.float32)
    _, res = _engine_run(X, 10000, 0.01, 4)
    ref = oracle.gauss_skeleton(oracle.corrcoef(X.astype(np.float64)), 10000, 0.01, 4)
    _compare(res, ref, f"C2 seed {seed}", pmax_atol=5e-5)
    for l in range(5):
        assert res.stats[l]["useful"] == ref["executed"][l]


def test_high_levels_match_oracle():
    """Levels 5-8, the deepest kernels built: a dense random SEM whose skeleton
    still has tests at level 7, and an equicorrelated block (every correlation
    0.5, n = 100 000) whose partial correlations stay far above the threshold,
    so that every subset of every pair runs up to level 8.  Skeleton, sepsets,
    pmax and the useful counts against the FP64 oracle."""
    d = syn.gaussian_sem(100, 10000, 8.0, seed=11)
    _, res = _engine_run(d.data, 10000, 0.01, 8)
    ref = oracle.gauss_skeleton(oracle.corrcoef(d.data), 10000, 0.01, 8)
    assert ref["executed"][7] > 0
    _compare(res, ref, "dense SEM, max level 8", pmax_atol=5e-5)  # FP32 z at n = 10 000, as for C2
    for l in range(9):
        assert res.stats[l]["useful"] == ref["executed"][l]
    p, n = 12, 100000
    C = np.full((p, p), 0.5)
    np.fill_diagonal(C, 1.0)
    _, res = _engine_run(None, n, 0.01, 8, C=C)
    ref = oracle.gauss_skeleton(C, n, 0.01, 8)
    assert ref["executed"][8] == 66 * 2 * 45  # both sides of every pair, C(10, 8) subsets
    _compare(res, ref, "equicorrelated, max level 8")
    for l in range(9):
        assert res.stats[l]["useful"] == ref["executed"][l]


def _mediators(n, k, seed):
    """i -> m_1..m_k -> j: the edge i - j survives every level below k and
    goes at level k with separation set {m_1..m_k}."""
    rng = np.random.default_rng(seed)
    i = rng.standard_normal(n)
    m = 0.6 * i[:, None] + rng.standard_normal((n, k))
    j = 0.4 * m.sum(1) + rng.standard_normal(n)
    return np.column_stack([i, m, j])


def test_levels_above_8_match_oracle():
    """Levels 9-12 (the kernel that takes the level at run time): ten
    mediators between two variables, so the edge goes at level 10 with all ten
    as its separation set; and an equicorrelated block of 14 variables, whose
    every subset of every pair runs up to level 12.  Skeleton, sepsets, pmax
    and the per-level test counts against the FP64 oracle."""
    X = _mediators(100000, 10, 5)
    _, res = _engine_run(X, 100000, 0.01, 10)
    ref = oracle.gauss_skeleton(oracle.corrcoef(X), 100000, 0.01, 10)
    assert ref["executed"][9] > 0 and ref["executed"][10] > 0
    assert ref["skeleton"][0, 11] == 0 and list(ref["sepsets"][0, 11]) == list(range(1, 11))
    _compare(res, ref, "ten mediators, max level 10", pmax_atol=5e-5)
    for l in range(11):
        assert res.stats[l]["useful"] == ref["executed"][l], l
    p, n = 14, 100000
    C = np.full((p, p), 0.5)
    np.fill_diagonal(C, 1.0)
    _, res = _engine_run(None, n, 0.01, 12, C=C)
    ref = oracle.gauss_skeleton(C, n, 0.01, 12)
    assert ref["executed"][12] == 91 * 2     # both sides of every pair, C(12, 12) subsets
    _compare(res, ref, "equicorrelated, max level 12")
    for l in range(13):
        assert res.stats[l]["useful"] == ref["executed"][l], l
    iu = np.triu_indices(p, 1)
    # levels 9-12 decide in FP64 with the reference's operation order
    assert np.abs(res.zmin[iu].astype(np.float64) - ref["zmin"][iu]).max() <= 1e-6


def test_levels_above_8_batched_and_captured():
    """A job that runs to level 10 in a batch with shallow jobs, launched
    eagerly and replayed from a captured graph: the same bits as alone."""
    from gpucsl_amd.batch import GaussJobState, GaussianBatch, run_plain

    X = _mediators(50000, 10, 6).astype(np.float32)
    Y = syn.gaussian_sem(80, 3000, 3.0, seed=8).data.astype(np.float32)
    jobs, plain = [], []
    for lst in (jobs, plain):
        a = GaussJobState(50000, 12, 10, alpha=0.01)
        a.set_data(X)
        b = GaussJobState(3000, 80, 3, alpha=0.01)
        b.set_data(Y)
        lst += [a, b]
    run_plain(plain)
    bt = GaussianBatch(jobs)
    bt.capture()
    for rep in range(3):
        bt.run()
        torch.cuda.synchronize()
        for a, c in zip(bt.results(), [j.results() for j in plain]):
            assert (a.skeleton == c.skeleton).all() and (a.separation_sets == c.separation_sets).all()
            assert np.array_equal(a.zmin.view(np.uint32), c.zmin.view(np.uint32))
    r = bt.results()[0]
    assert r.skeleton[0, 11] == 0 and list(r.separation_sets[0, 11]) == list(range(1, 11))


@pytest.mark.parametrize("seed", range(5))
def test_integer_gram_parity(seed):
    """float32 data with p = 400 (the integer-digit Gram path): skeleton,
    sepsets and the useful test counts identical to the FP64 oracle on the
    same float32 values, over five seeds."""
    n = 4000
    X = syn.gaussian_sem(400, n, 4.0, seed=500 + seed).data.astype(np.float32)
    _, res = _engine_run(X, n, 0.01, 3)
    ref = oracle.gauss_skeleton(oracle.corrcoef(X.astype(np.float64)), n, 0.01, 3)
    _compare(res, ref, f"integer Gram seed {seed}", pmax_atol=5e-5)
    for l in range(4):
        assert res.stats[l]["useful"] == ref["executed"][l]


def test_more_variables_than_samples():
    """p = 300 float32 columns (integer Gram path) from n = 150 rows: C has
    rank n - 1, and the skeleton, sepsets and useful test counts still equal
    the FP64 oracle's on the same values."""
    n = 150
    X = syn.gaussian_sem(300, n, 3.0, seed=77).data.astype(np.float32)
    _, res = _engine_run(X, n, 0.01, 3)
    ref = oracle.gauss_skeleton(oracle.corrcoef(X.astype(np.float64)), n, 0.01, 3)
    _compare(res, ref, "p > n", pmax_atol=5e-5)
    for l in range(4):
        assert res.stats[l]["useful"] == ref["executed"][l]


def test_float64_data_above_256_columns():
    """float64 data with p = 300 (the FP64 Gram on 64 x 64 tiles): skeleton,
    sepsets and useful test counts equal the FP64 oracle's."""
    n = 2500
    X = syn.gaussian_sem(300, n, 3.0, seed=303).data
    _, res = _engine_run(X, n, 0.01, 3)
    ref = oracle.gauss_skeleton(oracle.corrcoef(X), n, 0.01, 3)
    _compare(res, ref, "float64 p=300", pmax_atol=5e-5)
    for l in range(4):
        assert res.stats[l]["useful"] == ref["executed"][l]


def test_few_samples_zero_thresholds_at_high_levels():
    """n = 9 leaves nonzero thresholds |Phi^-1(a/2)| / sqrt(n - l - 4) only for
    levels 0..3 (min(50, n - 3) - 2 of them); from level 4 the threshold is 0
    and no test removes an edge.  An equicorrelated C (rho = 0.9, p = 8) at
    alpha = 0.9 keeps every edge, so every level 1..6 runs all of its tests;
    at alpha = 0.5 every edge goes at level 2.  Skeleton, sepsets, pmax and
    per-level counts against the oracle."""
    n, p = 9, 8
    C = np.full((p, p), 0.9)
    np.fill_diagonal(C, 1.0)
    for alpha, edges in ((0.9, 28), (0.5, 0)):
        _, res = _engine_run(None, n, alpha, 6, C=C)
        ref = oracle.gauss_skeleton(C, n, alpha, 6)
        _compare(res, ref, f"n=9 alpha={alpha}", pmax_atol=5e-5)
        assert int(res.skeleton.sum()) // 2 == edges
        for l in range(7):
            assert res.stats[l]["useful"] == ref["executed"][l]


def test_user_correlation_matrix_path():
    d = syn.gaussian_sem(120, 1500, 3.0, seed=21)
    C = oracle.corrcoef(d.data)
    _, res = _engine_run(None, 1500, 0.05, 3, C=C)
    ref = oracle.gauss_skeleton(C, 1500, 0.05, 3)
    _compare(res, ref, "user C")


def test_tight_margin_forces_no_recheck_paths_and_huge_margin_rechecks_all():
    d = syn.gaussian_sem(60, 800, 3.0, seed=3)
    ref = oracle.gauss_skeleton(oracle.corrcoef(d.data), 800, 0.05, 3)
    _, res_all = _engine_run(d.data, 800, 0.05, 3, margin=1e30)
    _compare(res_all, ref, "recheck everything", pmax_atol=1e-6)
    for l in range(1, 4):
        assert res_all.stats[l]["rechecked"] == res_all.stats[l]["executed"]


def test_deterministic_repeat():
    d = syn.gaussian_sem(200, 3000, 4.0, seed=5)
    eng, r1 = _engine_run(d.data, 3000, 0.01, 4)
    eng.run()
    r2 = eng.results()
    assert (r1.skeleton == r2.skeleton).all()
    assert (r1.separation_sets == r2.separation_sets).all()
    assert (r1.pmax.view(np.uint32) == r2.pmax.view(np.uint32)).all()


def test_edge_cases_singular_and_constant():
    """Duplicated columns make conditioning sets exactly singular; the FP64
    reference arithmetic is then decided by the last bits of C, so both sides
    get the same correlation matrix.  A constant column gives NaN correlations,
    which never remove an edge (gpucsl/pc/cuda/gaussian_ci.cu:39, :111)."""
    n = 1500
    X = syn.gaussian_sem(12, n, 3.0, seed=9).data
    X = np.concatenate([X, X[:, :2], np.ones((n, 1))], axis=1)   # duplicates + constant column
    with np.errstate(all="ignore"):
        C = oracle.corrcoef(X)
    _, res = _engine_run(None, n, 0.01, 3, C=C)
    ref = oracle.gauss_skeleton(C, n, 0.01, 3)
    _compare(res, ref, "singular/constant", pmax_atol=1e-5)
    assert res.skeleton[14, :14].all()          # the NaN column keeps every edge


def test_max_level_zero_and_tiny():
    d = syn.gaussian_sem(30, 500, 3.0, seed=2)
    _, res = _engine_run(d.data, 500, 0.01, 0)
    ref = oracle.gauss_skeleton(oracle.corrcoef(d.data), 500, 0.01, 0)
    assert res.separation_sets.shape == (30, 30, 0)
    assert (res.skeleton == ref["skeleton"]).all()
    X2 = d.data[:, :2]
    _, r2 = _engine_run(X2, 500, 0.01, 0)
    assert r2.skeleton.shape == (2, 2)


def test_dropin_api_shapes():
    from gpucsl_amd.pc.discover_skeleton_gaussian import discover_skeleton_gpu_gaussian
    from gpucsl_amd.pc.helpers import init_pc_graph

    d = syn.gaussian_sem(40, 1000, 3.0, seed=4)
    out = discover_skeleton_gpu_gaussian(init_pc_graph(d.data), d.data, None, 0.05, 3, 40, 1000)
    res, runtime = out
    assert runtime > 0 and res.discover_skeleton_kernel_time > 0
    assert res.skeleton.dtype == np.uint16 and res.skeleton.shape == (40, 40)
    assert res.seperation_sets.shape == (40, 40, 3) and res.seperation_sets.dtype == np.int32
    assert res.pmax.dtype == np.float32
    ref = oracle.gauss_skeleton(oracle.corrcoef(d.data), 1000, 0.05, 3)
    assert (res.skeleton == ref["skeleton"]).all()
    assert (res.seperation_sets == ref["sepsets"]).all()


def test_dropin_runs_in_a_row_with_changing_shapes():
    """The drop-in skeleton call several times in one process, with the
    dataset's size changing between calls (p = 40, 300, 40 again): each call
    equals the oracle, so no state of an earlier call (workspaces, launch
    tables) leaks into the next."""
    from gpucsl_amd.pc.discover_skeleton_gaussian import discover_skeleton_gpu_gaussian
    from gpucsl_amd.pc.helpers import init_pc_graph

    for p, n, seed in [(40, 1000, 41), (300, 3000, 42), (40, 1000, 41)]:
        d = syn.gaussian_sem(p, n, 3.0, seed=seed)
        res, _ = discover_skeleton_gpu_gaussian(init_pc_graph(d.data), d.data, None, 0.05, 3, p, n)
        ref = oracle.gauss_skeleton(oracle.corrcoef(d.data), n, 0.05, 3)
        say(f"[drop-in p={p} n={n}] edges={int(res.skeleton.sum()) // 2}")
        assert (res.skeleton == ref["skeleton"]).all()
        assert (res.seperation_sets == ref["sepsets"]).all()


def test_single_dataset_graph_replay_matches_eager():
    """The whole single-dataset skeleton captured as one HIP graph: every
    replay, with other GPU work between replays, equals the eager run."""
    from gpucsl_amd.gaussian import GaussianEngine
    from gpucsl_amd.pc.helpers import get_gaussian_thresholds

    d = syn.gaussian_sem(120, 3000, 3.0, seed=77).data.astype(np.float32)
    X = torch.from_numpy(d).cuda()
    eng = GaussianEngine(120, 3000, 3, get_gaussian_thresholds(d, 0.01))
    eng.set_data(X)
    eng.run()
    torch.cuda.synchronize()
    want = eng.results()
    eng.capture(X)
    for rep in range(3):
        junk = (torch.randn(4096, device="cuda") * 2).sum()
        eng.replay()
        torch.cuda.synchronize()
        got = eng.results()
        assert (got.skeleton == want.skeleton).all() and (got.separation_sets == want.separation_sets).all()
        assert np.array_equal(got.pmax.view(np.uint32), want.pmax.view(np.uint32))
        assert [s["useful"] for s in got.stats] == [s["useful"] for s in want.stats]
        del junk


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_pcalg_mode_matches_oracle(seed):
    """pcalg conventions (thresholds sqrt(n-l-3), pMax over every test with its
    own scale, one-way sepsets) against the oracle's pcalg mode."""
    from gpucsl_amd.pc.discover_skeleton_gaussian import discover_skeleton_gpu_gaussian
    from gpucsl_amd.pc.helpers import init_pc_graph

    X = syn.gaussian_sem(50, 2000, 3.0, seed=seed).data.astype(np.float32)
    res, _ = discover_skeleton_gpu_gaussian(init_pc_graph(X), X, None, 0.01, 3, 50, 2000, pcalg_compat=True)
    ref = oracle.gauss_skeleton(oracle.corrcoef(X.astype(np.float64)), 2000, 0.01, 3, pcalg_compat=True)
    assert (res.skeleton == ref["skeleton"]).all()
    assert (res.seperation_sets == ref["sepsets"]).all()
    one_way = [(a, b) for a in range(50) for b in range(50)
               if ref["sepsets"][a, b, 0] != -1 and ref["sepsets"][b, a, 0] == -1]
    iu = np.triu_indices(50, 1)
    dp = np.abs(res.pmax[iu].astype(np.float64) - ref["pmax"][iu]).max()
    say(f"[pcalg seed {seed}] edges {int(res.skeleton.sum()) // 2}, one-way sepsets {len(one_way)}, "
        f"max|dpmax| {dp:.3g}")
    assert dp <= 2e-5
    plain = oracle.gauss_skeleton(oracle.corrcoef(X.astype(np.float64)), 2000, 0.01, 3)
    assert (np.abs(plain["pmax"][iu] - ref["pmax"][iu]).max() > 0) or len(one_way) > 0


@pytest.mark.parametrize("seed", range(10))
def test_pcalg_mode_fp64_c1(seed):
    """pcalg mode on C1's float64 data (p=50, n=2000, degree 3, L=3): the
    correlation is computed from the float64 values, so skeleton and sepsets
    equal the FP64 oracle's pcalg mode on the same data."""
    from gpucsl_amd.pc.discover_skeleton_gaussian import discover_skeleton_gpu_gaussian
    from gpucsl_amd.pc.helpers import init_pc_graph

    X = syn.gaussian_sem(50, 2000, 3.0, seed=seed).data
    assert X.dtype == np.float64
    res, _ = discover_skeleton_gpu_gaussian(init_pc_graph(X), X, None, 0.01, 3, 50, 2000, pcalg_compat=True)
    ref = oracle.gauss_skeleton(oracle.corrcoef(X), 2000, 0.01, 3, pcalg_compat=True)
    assert (res.skeleton == ref["skeleton"]).all()
    assert (res.seperation_sets == ref["sepsets"]).all()
    iu = np.triu_indices(50, 1)
    dp = np.abs(res.pmax[iu].astype(np.float64) - ref["pmax"][iu]).max()
    say(f"[pcalg fp64 C1 seed {seed}] edges {int(res.skeleton.sum()) // 2}, max|dpmax| {dp:.3g}")
    assert dp <= 2e-5


def test_pcalg_mode_deep_levels():
    """pcalg conventions up to level 8 on a dense SEM (degree 8, tests up to
    level 7): each level's own sqrt(n - l - 3) in the thresholds and in pMax."""
    from gpucsl_amd.pc.discover_skeleton_gaussian import discover_skeleton_gpu_gaussian
    from gpucsl_amd.pc.helpers import init_pc_graph

    p, n = 100, 10000
    X = syn.gaussian_sem(p, n, 8.0, seed=11).data.astype(np.float32)
    res, _ = discover_skeleton_gpu_gaussian(init_pc_graph(X), X, None, 0.01, 8, p, n, pcalg_compat=True)
    ref = oracle.gauss_skeleton(oracle.corrcoef(X.astype(np.float64)), n, 0.01, 8, pcalg_compat=True)
    assert ref["executed"][7] > 0
    assert (res.skeleton == ref["skeleton"]).all()
    assert (res.seperation_sets == ref["sepsets"]).all()
    iu = np.triu_indices(p, 1)
    dp = np.abs(res.pmax[iu].astype(np.float64) - ref["pmax"][iu]).max()
    say(f"[pcalg deep] edges {int(res.skeleton.sum()) // 2}, max|dpmax| {dp:.3g}")
    assert dp <= 5e-5  # FP32 z at n = 10 000, as for C2


def test_level_by_level_and_batch_of_one_equal_whole_skeleton():
    """The three ways of launching one dataset's levels give the same bits:
    the whole-skeleton entry point (level planning inside the compaction
    launch), level by level (gpucsl_gauss_init/level0/level), and a batch of
    one job (separate planning launch)."""
    from gpucsl_amd.batch import GaussianBatch, GaussJobState
    from gpucsl_amd.gaussian import GaussianEngine
    from gpucsl_amd.pc.helpers import get_gaussian_thresholds

    d = syn.gaussian_sem(150, 4000, 3.0, seed=5).data.astype(np.float32)
    X = torch.from_numpy(d).cuda()
    thr = get_gaussian_thresholds(d, 0.01)
    eng = GaussianEngine(150, 4000, 4, thr)
    eng.set_data(X)
    eng.run()
    torch.cuda.synchronize()
    want = eng.results()
    for rep in range(2):
        eng.run_levels()
        torch.cuda.synchronize()
        got = eng.results()
        say(f"[levels rep {rep}] edges={int(got.skeleton.sum()) // 2}")
        assert (got.skeleton == want.skeleton).all() and (got.separation_sets == want.separation_sets).all()
        assert np.array_equal(got.pmax.view(np.uint32), want.pmax.view(np.uint32))
    job = GaussJobState(4000, 150, 4, thresholds=thr)
    job.set_data(X)
    b = GaussianBatch([job])
    b.run()
    torch.cuda.synchronize()
    r = b.results()[0]
    assert (r.skeleton == want.skeleton).all() and (r.separation_sets == want.separation_sets).all()
    assert np.array_equal(r.pmax.view(np.uint32), want.pmax.view(np.uint32))


def test_fused_outputs_equal_separate_entry_points():
    """gpucsl_gauss_outputs equals gpucsl_pmax_from_zmin + gpucsl_skeleton_to_u16."""
    eng, res = _engine_run(syn.gaussian_sem(70, 1500, 3.0, seed=9).data, 1500, 0.05, 3)
    p, s = 70, nat.stream_handle()
    pm = torch.empty((p, p), dtype=torch.float32, device="cuda")
    sk = torch.empty((p, p), dtype=torch.int16, device="cuda")
    nat.call("gpucsl_pmax_from_zmin", eng.zmin.data_ptr(), p, eng.n, pm.data_ptr(), s)
    nat.call("gpucsl_skeleton_to_u16", eng.G.data_ptr(), p, sk.data_ptr(), s)
    torch.cuda.synchronize()
    assert np.array_equal(pm.cpu().numpy().view(np.uint32), res.pmax.view(np.uint32))
    assert np.array_equal(sk.cpu().numpy().view(np.uint16), res.skeleton)


def test_debug_build_runs_and_matches_release():
    """is_debug=True runs the index-checked debug build of the library (every
    check holds on valid input) and gives the release build's results."""
    from gpucsl_amd.pc.pc import DiscretePC, GaussianPC

    X = syn.gaussian_sem(120, 3000, 3.0, seed=41).data
    rel = GaussianPC(X, 3, 0.01).set_distribution_specific_options().execute().result
    dbg = GaussianPC(X, 3, 0.01, is_debug=True).set_distribution_specific_options().execute().result
    assert "debug" in nat._libs and not nat._debug
    assert set(rel.directed_graph.edges) == set(dbg.directed_graph.edges)
    assert (rel.separation_sets == dbg.separation_sets).all()
    assert np.array_equal(rel.pmax.view(np.uint32), dbg.pmax.view(np.uint32))
    d = syn.discrete_bn(25, 4000, 2.0, seed=42).data
    rel = DiscretePC(d, 3, 0.05).set_distribution_specific_options().execute().result
    dbg = DiscretePC(d, 3, 0.05, is_debug=True).set_distribution_specific_options().execute().result
    assert set(rel.directed_graph.edges) == set(dbg.directed_graph.edges)
    assert (rel.separation_sets == dbg.separation_sets).all()
    assert np.array_equal(rel.pmax.view(np.uint32), dbg.pmax.view(np.uint32))
