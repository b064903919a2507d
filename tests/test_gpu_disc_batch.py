"""Batched discrete skeletons (gpucsl_disc_batch_*) on the GPU: every job of a
batch equals the same job run alone through the single-dataset entry points,
bit for bit (skeleton, separation sets, pmax and the per-level statistics);
two jobs also against the FP64 oracle; more jobs than one launch holds; level
by level equals the whole; a captured graph replays to the eager result.  GPU
box only."""
import sys

import numpy as np
import pytest

import oracle
from gpucsl_amd import synthetic as syn

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

ALPHA = 0.05

# p, n, card_low, card_high, max_level, seed
SPECS = [
    (30, 4000, 2, 5, 3, 1),
    (12, 2500, 17, 20, 2, 2),   # more than 16 categories: level-0 tables counted pair by pair
    (45, 1500, 2, 4, 2, 3),
    (20, 6000, 2, 3, 0, 4),     # level 0 only
    (8, 3000, 2, 2, 3, 5),
    (60, 2000, 2, 5, 3, 6),
    (14, 2500, 14, 16, 1, 7),   # 16 categories: level 0 from the one-hot Gram matrix
]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _deep_codes():
    """A binary variable, ten noisy copies and a threshold of their sum: the
    edge between the first and the last goes at level 10 (the kernels that
    take the level at run time)."""
    rng = np.random.default_rng(5)
    n, k = 50000, 10
    i = rng.integers(0, 2, n)
    m = (i[:, None] ^ (rng.random((n, k)) < 0.2)).astype(np.int64)
    j = ((m.sum(1) + rng.integers(0, 2, n)) > 5).astype(np.int64)
    return np.column_stack([i, m, j]).astype(np.uint8)


def _datasets(specs):
    out = []
    for (p, n, lo, hi, L, seed) in specs:
        out.append((syn.discrete_bn(p, n, 2.0, seed=seed, card_low=lo, card_high=hi).data, L))
    return out


def _engine(codes, L):
    from gpucsl_amd.discrete import DiscreteEngine

    codes = np.asarray(codes, np.uint8)
    n, p = codes.shape
    dims = (codes.max(axis=0).astype(np.int64) + 1).astype(np.uint8)
    eng = DiscreteEngine(p, n, L, ALPHA, int(dims.max()))
    eng.set_codes(codes, dims)
    return eng


def _alone(datasets):
    res = []
    for codes, L in datasets:
        e = _engine(codes, L)
        e.run()
        res.append(e.results())
    return res


def _batch(datasets):
    from gpucsl_amd.discrete import DiscreteBatch

    return DiscreteBatch([_engine(codes, L) for codes, L in datasets])


def _same(a, b, label):
    ok = ((a.skeleton == b.skeleton).all() and (a.separation_sets == b.separation_sets).all()
          and np.array_equal(a.pmax.view(np.uint32), b.pmax.view(np.uint32)) and a.stats == b.stats)
    say(f"[{label}] identical={ok} edges={int(a.skeleton.sum()) // 2}")
    assert ok, label


def test_disc_batch_equals_alone_and_oracle():
    data = _datasets(SPECS) + [(_deep_codes(), 10)]
    want = _alone(data)
    b = _batch(data)
    b.launch()
    torch.cuda.synchronize()
    got = b.results()
    for k, (g, w) in enumerate(zip(got, want)):
        _same(g, w, f"job {k} p={data[k][0].shape[1]} n={data[k][0].shape[0]} L={data[k][1]}")
    for k in (0, len(data) - 1):
        codes, L = data[k]
        dims = (codes.max(axis=0).astype(np.int64) + 1).astype(np.uint8)
        ref = oracle.disc_skeleton(np.ascontiguousarray(codes.T), dims, ALPHA, L)
        iu = np.triu_indices(codes.shape[1], 1)
        assert (got[k].skeleton == ref["skeleton"]).all() and (got[k].separation_sets == ref["sepsets"]).all()
        assert np.abs(got[k].pmax[iu].astype(np.float64) - ref["pmax"][iu]).max() <= 1e-6
        assert [s["executed"] for s in got[k].stats] == [int(x) for x in ref["executed"][: L + 1]]
    assert got[-1].skeleton[0, 11] == 0


def test_disc_batch_more_jobs_than_one_launch():
    """21 jobs: the batched entry points run them in launches of
    GPUCSL_DISC_BATCH_JOBS; every job still equals its run alone."""
    from gpucsl_amd.discrete import BATCH_JOBS

    rng = np.random.default_rng(11)
    specs = [(int(rng.integers(6, 26)), int(rng.integers(300, 3001)), 2, int(rng.integers(2, 6)),
              int(rng.integers(1, 4)), 100 + k) for k in range(21)]
    assert len(specs) > BATCH_JOBS
    data = _datasets(specs)
    want = _alone(data)
    b = _batch(data)
    b.launch()
    torch.cuda.synchronize()
    for k, (g, w) in enumerate(zip(b.results(), want)):
        _same(g, w, f"job {k}")


def test_disc_batch_level_by_level_equals_whole():
    data = _datasets(SPECS[:4])
    want = _alone(data)
    b = _batch(data)
    b.start()
    for l in range(1, b.max_level + 1):
        b.level(l)
    b.finish()
    torch.cuda.synchronize()
    for k, (g, w) in enumerate(zip(b.results(), want)):
        _same(g, w, f"level by level, job {k}")


def test_disc_batch_graph_replay_matches_eager():
    data = _datasets(SPECS[:5])
    want = _alone(data)
    b = _batch(data)
    b.capture()
    for rep in range(3):
        b.run()
        torch.cuda.synchronize()
        for k, (g, w) in enumerate(zip(b.results(), want)):
            _same(g, w, f"replay {rep}, job {k}")
        _alone(data[:1])   # eager work of the program between replays


def test_disc_batch_edge_shapes_and_restricted_workspace():
    """A two-variable job, an n that is not a multiple of 16, and a job whose
    tables exceed the LDS (8 categories at level 3) with a workspace of only a
    few global table slices, which caps the workgroups of every launch it is
    in: each equals its run alone on the same workspace size."""
    from gpucsl_amd.discrete import DiscreteBatch, DiscreteEngine, disc_workspace

    d_big = syn.discrete_bn(10, 3000, 3.0, seed=41, card_low=8, card_high=8).data
    dims_big = (d_big.max(axis=0).astype(np.int64) + 1).astype(np.uint8)
    from gpucsl_amd import _native as nat

    full = disc_workspace(10, 3, int(dims_big.max()))
    least = int(nat.load().gpucsl_disc_min_workspace_bytes(10, 3, int(dims_big.max())))
    assert full > least
    limit = least + 3 * (full - least) // 1024   # about three table slices
    cases = [(syn.discrete_bn(2, 500, 1.0, seed=42).data, 2, None),
             (syn.discrete_bn(15, 1001, 2.0, seed=43).data, 3, None),
             (d_big, 3, limit),
             (syn.discrete_bn(25, 2000, 2.0, seed=44).data, 2, None)]

    def engine(codes, L, ws_limit):
        dims = (codes.max(axis=0).astype(np.int64) + 1).astype(np.uint8)
        e = DiscreteEngine(codes.shape[1], codes.shape[0], L, ALPHA, int(dims.max()), ws_limit=ws_limit)
        e.set_codes(codes, dims)
        return e

    want = []
    for codes, L, lim in cases:
        e = engine(codes, L, lim)
        e.run()
        want.append(e.results())
    b = DiscreteBatch([engine(codes, L, lim) for codes, L, lim in cases])
    b.launch()
    torch.cuda.synchronize()
    for k, (g, w) in enumerate(zip(b.results(), want)):
        _same(g, w, f"edge case {k}")
