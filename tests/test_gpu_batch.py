"""Batched Gaussian skeletons (many datasets in shared launches) on the GPU:
every job of a batch equals the same job run alone through the single-dataset
path (bit for bit) and the FP64 oracle (skeleton and separation sets); a
captured graph replays to the eager result.  GPU box only."""
import sys

import numpy as np
import pytest

import oracle
from gpucsl_amd import synthetic as syn

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _jobs(specs):
    from gpucsl_amd.batch import GaussJobState

    jobs, data = [], []
    for (p, n, L, alpha, seed) in specs:
        X = syn.gaussian_sem(p, n, 3.0, seed=seed).data.astype(np.float32)
        j = GaussJobState(n, p, L, alpha=alpha)
        j.set_data(X)
        jobs.append(j)
        data.append(X)
    return jobs, data


def _plain_results(specs):
    from gpucsl_amd.batch import run_plain

    jobs, _ = _jobs(specs)
    run_plain(jobs)
    torch.cuda.synchronize()
    return [j.results() for j in jobs]


def _same(a, b, label):
    ok = ((a.skeleton == b.skeleton).all() and (a.separation_sets == b.separation_sets).all()
          and np.array_equal(a.zmin.view(np.uint32), b.zmin.view(np.uint32))
          and np.array_equal(a.pmax.view(np.uint32), b.pmax.view(np.uint32)))
    if not ok:
        say(f"[{label}] differs: skel {(a.skeleton != b.skeleton).sum()} sep "
            f"{(a.separation_sets != b.separation_sets).sum()}")
    return ok


# (p, n, max_level, alpha, seed); the p = 300 and p = 420 jobs take the integer Gram path
SPECS = [(20, 300, 3, 0.01, 1), (57, 2000, 3, 0.01, 2), (80, 1500, 2, 0.05, 3), (33, 5000, 0, 0.01, 4),
         (120, 3000, 3, 0.01, 5), (64, 999, 1, 0.01, 6), (200, 5000, 3, 0.01, 7), (45, 700, 3, 0.05, 8),
         (300, 3000, 3, 0.01, 9), (420, 2500, 2, 0.01, 10)]


def test_batch_equals_alone_and_oracle():
    from gpucsl_amd.batch import GaussianBatch

    say(f"batch of {len(SPECS)} jobs: {SPECS}")
    jobs, data = _jobs(SPECS)
    b = GaussianBatch(jobs)
    b.launch()
    torch.cuda.synchronize()
    got = b.results()
    alone = _plain_results(SPECS)
    refs = [oracle.gauss_skeleton(oracle.corrcoef(X.astype(np.float64)), n, alpha, L)
            for (p, n, L, alpha, seed), X in zip(SPECS, data)]
    for k, ((p, n, L, alpha, seed), ref) in enumerate(zip(SPECS, refs)):
        assert _same(got[k], alone[k], f"job {k} batch vs alone")
        assert (got[k].skeleton == ref["skeleton"]).all(), f"job {k} skeleton vs oracle"
        assert (got[k].separation_sets == ref["sepsets"]).all(), f"job {k} sepsets vs oracle"
        iu = np.triu_indices(p, 1)
        dp = np.abs(got[k].pmax[iu].astype(np.float64) - ref["pmax"][iu]).max()
        say(f"  job {k} p={p} n={n} L={L}: edges={int(got[k].skeleton.sum()) // 2} max|dpmax|={dp:.3g}")
        assert dp <= 2e-5
    st = b.batch_stats()
    oracle_exec = [sum(int(ref["executed"][l]) if l <= L else 0 for (p, n, L, a, s), ref in zip(SPECS, refs))
                   for l in range(4)]
    say("  batch useful per level", [d["useful"] for d in st.levels], "oracle", oracle_exec)
    assert [d["useful"] for d in st.levels] == oracle_exec


def test_batch_recheck_everything_equals_plain_and_oracle():
    """Margin 1e30 sends every level >= 1 test of every job through the FP64
    recheck queue (items from several jobs in one queue, pending side-A pairs
    continued on their side B): each job equals its plain run at the same
    margin bit for bit and the oracle's skeleton and sepsets, and every
    executed test was rechecked."""
    from gpucsl_amd.batch import GaussianBatch, run_plain

    specs = SPECS[:6]
    jobs, data = _jobs(specs)
    b = GaussianBatch(jobs, margin=1e30)
    b.launch()
    torch.cuda.synchronize()
    got = b.results()
    plain, _ = _jobs(specs)
    run_plain(plain, margin=1e30)
    torch.cuda.synchronize()
    for k, ((p, n, L, alpha, seed), X) in enumerate(zip(specs, data)):
        ref = oracle.gauss_skeleton(oracle.corrcoef(X.astype(np.float64)), n, alpha, L)
        assert _same(got[k], plain[k].results(), f"job {k} batch vs plain at margin 1e30")
        assert (got[k].skeleton == ref["skeleton"]).all(), f"job {k} skeleton vs oracle"
        assert (got[k].separation_sets == ref["sepsets"]).all(), f"job {k} sepsets vs oracle"
    st = b.batch_stats()
    for l in range(1, 4):
        say(f"  level {l}: {st.levels[l]}")
        assert st.levels[l]["rechecked"] == st.levels[l]["executed"]


def test_batch_with_deep_levels_equals_alone_and_oracle():
    """Jobs that stop at different levels up to 8 (dense SEMs, degree 8, with
    tests up to level 7) in one batch: each job bit-identical to its run alone,
    skeleton and sepsets equal to the FP64 oracle, and the batch's useful
    counts per level equal to the oracle's."""
    from gpucsl_amd.batch import GaussianBatch, GaussJobState, run_plain

    specs = [(100, 10000, 8, 11), (60, 5000, 6, 3), (70, 6000, 3, 9)]

    def make():
        jobs, data = [], []
        for (p, n, L, seed) in specs:
            X = syn.gaussian_sem(p, n, 8.0, seed=seed).data.astype(np.float32)
            j = GaussJobState(n, p, L, alpha=0.01)
            j.set_data(X)
            jobs.append(j)
            data.append(X)
        return jobs, data

    jobs, data = make()
    b = GaussianBatch(jobs)
    b.launch()
    torch.cuda.synchronize()
    got = b.results()
    alone, _ = make()
    run_plain(alone)
    torch.cuda.synchronize()
    refs = [oracle.gauss_skeleton(oracle.corrcoef(X.astype(np.float64)), n, 0.01, L)
            for (p, n, L, seed), X in zip(specs, data)]
    assert refs[0]["executed"][7] > 0
    for k, ((p, n, L, seed), ref) in enumerate(zip(specs, refs)):
        assert _same(got[k], alone[k].results(), f"deep job {k} batch vs alone")
        assert (got[k].skeleton == ref["skeleton"]).all(), f"deep job {k} skeleton vs oracle"
        assert (got[k].separation_sets == ref["sepsets"]).all(), f"deep job {k} sepsets vs oracle"
    st = b.batch_stats()
    oracle_exec = [sum(int(ref["executed"][l]) if l <= L else 0 for (p, n, L, s), ref in zip(specs, refs))
                   for l in range(9)]
    say("  deep batch useful per level", [d["useful"] for d in st.levels], "oracle", oracle_exec)
    assert [d["useful"] for d in st.levels] == oracle_exec


def test_batch_membership_does_not_change_results():
    """A job's result is the same bits in two different batches."""
    from gpucsl_amd.batch import GaussianBatch

    a_specs = SPECS[:3] + [(150, 4000, 3, 0.01, 11), (330, 3500, 3, 0.01, 12)]
    b_specs = [(330, 3500, 3, 0.01, 12), (150, 4000, 3, 0.01, 11)] + SPECS[5:]
    ja, _ = _jobs(a_specs)
    jb, _ = _jobs(b_specs)
    ba, bb = GaussianBatch(ja), GaussianBatch(jb)
    ba.launch()
    bb.launch()
    torch.cuda.synchronize()
    assert _same(ba.results()[3], bb.results()[1], "job in two batches")
    assert _same(ba.results()[4], bb.results()[0], "integer-Gram job in two batches")


def test_graph_replay_matches_eager():
    from gpucsl_amd.batch import GaussianBatch

    jobs, _ = _jobs(SPECS)
    b = GaussianBatch(jobs)
    b.launch()
    torch.cuda.synchronize()
    want = b.results()
    want_stats = b.batch_stats().levels
    b.capture()
    for rep in range(3):
        # eager work between replays
        junk = torch.randn(1000, device="cuda").sum()
        b.run()
        torch.cuda.synchronize()
        got = b.results()
        for k in range(len(SPECS)):
            assert _same(got[k], want[k], f"replay {rep} job {k}")
        assert b.batch_stats().levels == want_stats
        del junk


def test_bootstrap_batch_c4_slice():
    """Bootstrap resamples of the p=200, n=5000 dataset (seed 300, rows drawn
    with seed 300+b) plus uneven datasets, as in the batched workload."""
    from gpucsl_amd.batch import GaussJobState, GaussianBatch, run_plain

    base = syn.gaussian_sem(200, 5000, 3.0, seed=300).data
    jobs, plain, xs = [], [], []
    for b in range(6):
        X = syn.bootstrap(base, 300 + b).astype(np.float32)
        xs.append((X, 5000))
    rng = np.random.Generator(np.random.PCG64(4000))
    for k in range(3):
        p, n = int(rng.integers(50, 501)), int(rng.integers(1000, 20001))
        xs.append((syn.gaussian_sem(p, n, 3.0, seed=400 + k).data.astype(np.float32), n))
    for X, n in xs:
        for lst in (jobs, plain):
            j = GaussJobState(n, X.shape[1], 3, alpha=0.01)
            j.set_data(X)
            lst.append(j)
    bt = GaussianBatch(jobs)
    bt.launch()
    run_plain(plain)
    torch.cuda.synchronize()
    for k, (a, c) in enumerate(zip(bt.results(), [j.results() for j in plain])):
        assert _same(a, c, f"c4 job {k}")
    for k in (0, 7):
        X, n = xs[k]
        ref = oracle.gauss_skeleton(oracle.corrcoef(X.astype(np.float64)), n, 0.01, 3)
        r = bt.results()[k]
        assert (r.skeleton == ref["skeleton"]).all() and (r.separation_sets == ref["sepsets"]).all()
        say(f"  c4 job {k} p={X.shape[1]} n={n} matches oracle; edges={int(r.skeleton.sum()) // 2}")


def test_mixed_inputs_rejected():
    from gpucsl_amd.batch import GaussJobState, GaussianBatch
    from gpucsl_amd.cli.cli_util import GpucslError

    a = GaussJobState(500, 10, 1, alpha=0.01)
    a.set_data(np.random.default_rng(0).standard_normal((500, 10)).astype(np.float32))
    c = GaussJobState(500, 10, 1, alpha=0.01)
    c.set_correlation(np.eye(10))
    with pytest.raises(GpucslError):
        GaussianBatch([a, c]).launch()


def test_full_c4_batch_equals_alone_and_oracle():
    """The whole batched workload: 256 bootstrap resamples of the p=200,
    n=5000 dataset plus the 64 uneven datasets (p ~ U[50, 500], n ~ U[1000,
    20000]), max level 3, in one batch (the batched benchmark's jobs).  Every
    one of the 320 jobs equals the same job through the single-dataset path,
    bit for bit; five jobs, two of them above 256 columns (integer Gram path),
    equal the FP64 oracle's skeleton and separation sets."""
    import bench
    from gpucsl_amd.batch import GaussJobState, GaussianBatch, run_plain

    base, rows, sets = bench.batched_datasets(256, 64)
    base_dev = torch.from_numpy(base).cuda()
    jobs, plain, xs = [], [], []
    for r in rows:
        X = base_dev[torch.from_numpy(r).cuda()]
        xs.append((base, r))
        for lst in (jobs, plain):
            j = GaussJobState(5000, 200, 3, alpha=0.01)
            j.set_data(X)
            lst.append(j)
    for X in sets:
        n, p = X.shape
        xs.append((X, None))
        for lst in (jobs, plain):
            j = GaussJobState(n, p, 3, alpha=0.01)
            j.set_data(X)
            lst.append(j)
    torch.cuda.synchronize()
    say(f"c4: {len(jobs)} jobs, {sum(j.nbytes for j in jobs) / 2**30:.2f} GiB of job state per copy")
    bt = GaussianBatch(jobs)
    bt.launch()
    run_plain(plain)
    torch.cuda.synchronize()
    got = bt.results()
    for k, (a, j) in enumerate(zip(got, plain)):
        assert _same(a, j.results(), f"c4 job {k}"), f"job {k} differs from its run alone"
    big = [256 + k for k, X in enumerate(sets) if X.shape[1] > 256]
    assert len(big) >= 2
    for k in [0, 255, 256, big[0], big[1]]:
        X, r = xs[k]
        X64 = (X[r] if r is not None else X).astype(np.float64)
        n = X64.shape[0]
        ref = oracle.gauss_skeleton(oracle.corrcoef(X64), n, 0.01, 3)
        assert (got[k].skeleton == ref["skeleton"]).all() and (got[k].separation_sets == ref["sepsets"]).all(), k
        say(f"  c4 job {k} p={X64.shape[1]} n={n} equals the oracle; edges {int(got[k].skeleton.sum()) // 2}")
