"""Serving loop on the GPU: requests admitted by HBM footprint while others are
in flight give the same bits as the same request alone through the
single-dataset path, for Gaussian and discrete requests.  GPU box only."""
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


def _requests(seed, n_gauss, n_disc, small=True):
    from gpucsl_amd.serving import Request

    rng = np.random.Generator(np.random.PCG64(seed))
    reqs = []
    kinds = ["gaussian"] * n_gauss + ["discrete"] * n_disc
    rng.shuffle(kinds)
    for k, kind in enumerate(kinds):
        if kind == "gaussian":
            p, n = int(rng.integers(20, 151)), int(rng.integers(500, 3001))
            X = syn.gaussian_sem_device(p, n, 3.0, seed=seed * 100 + k)
            reqs.append(Request("gaussian", n, p, max_level=int(rng.integers(1, 4)), alpha=0.01, X=X))
        else:
            p, n = int(rng.integers(10, 41)), int(rng.integers(1000, 5001))
            codes, dims = syn.discrete_bn_device(p, n, 2.0, seed=seed * 100 + k)
            reqs.append(Request("discrete", n, p, max_level=3, alpha=0.01, codes=codes, dims=dims))
    torch.cuda.synchronize()
    return reqs


def _alone(req):
    from gpucsl_amd.serving import run_alone

    st = run_alone(req)
    torch.cuda.synchronize()
    return st.results()


def _same(served, alone, kind):
    ok = (served.skeleton == alone.skeleton).all() and (served.separation_sets == alone.separation_sets).all()
    ok = ok and np.array_equal(served.pmax.view(np.uint32), alone.pmax.view(np.uint32))
    if kind == "gaussian":
        ok = ok and np.array_equal(served.zmin.view(np.uint32), alone.zmin.view(np.uint32))
    return ok


def test_served_equals_alone_under_continuous_batching():
    from gpucsl_amd.serving import SkeletonServer

    reqs = _requests(7, 10, 5)
    srv = SkeletonServer()
    done = []
    # staggered arrivals: a few requests, a few ticks, more requests
    for k, r in enumerate(reqs):
        srv.submit(r)
        if k % 3 == 2:
            done += srv.tick()
            done += srv.tick()
    done += srv.drain()
    assert len(done) == len(reqs) and srv.idle
    say(f"served {len(done)} requests in {srv.ticks} ticks, largest Gaussian batch {srv.max_batch}")
    assert srv.max_batch >= 2
    by_id = {c.rid: c for c in done}
    for r in reqs:
        c = by_id[r.rid]
        assert c.latency > 0
        assert _same(c.result, _alone(r), r.kind), f"request {r.rid} ({r.kind} p={r.p} n={r.n}) differs"


def test_deep_level_requests_served_equal_alone():
    """Requests that run to levels 6-8 (dense Gaussian SEMs, a binary network)
    served together with shallow ones: each result bit-identical to its run
    alone."""
    from gpucsl_amd.serving import Request, SkeletonServer

    reqs = [Request("gaussian", 10000, 100, max_level=8, alpha=0.01,
                    X=syn.gaussian_sem_device(100, 10000, 8.0, seed=11)),
            Request("gaussian", 5000, 60, max_level=6, alpha=0.01,
                    X=syn.gaussian_sem_device(60, 5000, 8.0, seed=3))]
    codes, dims = syn.discrete_bn_device(16, 50000, 5.0, seed=2, card_low=2, card_high=2)
    reqs.append(Request("discrete", 50000, 16, max_level=8, alpha=0.05, codes=codes, dims=dims))
    reqs += _requests(9, 3, 2)
    torch.cuda.synchronize()
    srv = SkeletonServer()
    for r in reqs:
        srv.submit(r)
    done = srv.drain()
    assert len(done) == len(reqs) and srv.idle
    by_id = {c.rid: c for c in done}
    for r in reqs:
        assert _same(by_id[r.rid].result, _alone(r), r.kind), f"request {r.rid} ({r.kind} p={r.p} L={r.max_level}) differs"
    say(f"deep-level mix: {len(reqs)} requests in {srv.ticks} ticks, all equal alone")


def test_results_on_device_equal_alone_and_hold_their_range():
    """results_on_device: results are device tensors of the job's own range,
    the same bits as the request alone, and the range stays in use until the
    caller releases the result."""
    from gpucsl_amd.serving import SkeletonServer

    reqs = _requests(9, 6, 3)
    srv = SkeletonServer(results_on_device=True)
    for r in reqs:
        srv.submit(r)
    done = srv.drain()
    assert len(done) == len(reqs)
    held = srv.used
    assert held > 0                      # every finished job still holds its device range
    for c in done:
        r = next(x for x in reqs if x.rid == c.rid)
        res = c.result
        assert res.skeleton.is_cuda and res.pmax.is_cuda
        a = _alone(r)
        assert (res.skeleton.cpu().numpy().view(np.uint16) == a.skeleton).all()
        assert (res.separation_sets.cpu().numpy() == a.separation_sets).all()
        assert np.array_equal(res.pmax.cpu().numpy().view(np.uint32), a.pmax.view(np.uint32))
        if r.kind == "gaussian":
            assert np.array_equal(res.zmin.cpu().numpy().view(np.uint32), a.zmin.view(np.uint32))
        c.release()
    assert srv.used == 0


def test_served_matches_oracle_small():
    """Served Gaussian and discrete results against the FP64 oracle."""
    from gpucsl_amd.serving import Request, SkeletonServer

    d = syn.gaussian_sem(40, 2000, 3.0, seed=3)
    X = d.data.astype(np.float32)
    b = syn.discrete_bn(15, 3000, 2.0, seed=4)
    dims = (b.data.max(axis=0) + 1).astype(np.uint8)
    codes = torch.from_numpy(np.ascontiguousarray(b.data.T)).cuda()
    srv = SkeletonServer()
    srv.submit(Request("gaussian", 2000, 40, 3, 0.01, X=torch.from_numpy(X).cuda()))
    srv.submit(Request("discrete", 3000, 15, 2, 0.05, codes=codes, dims=torch.from_numpy(dims).cuda()))
    out = {c.kind: c for c in srv.drain()}
    ref = oracle.gauss_skeleton(oracle.corrcoef(X.astype(np.float64)), 2000, 0.01, 3)
    g = out["gaussian"].result
    assert (g.skeleton == ref["skeleton"]).all() and (g.separation_sets == ref["sepsets"]).all()
    rd = oracle.disc_skeleton(np.ascontiguousarray(b.data.T), dims, 0.05, 2)
    dd = out["discrete"].result
    assert (dd.skeleton == rd["skeleton"]).all() and (dd.separation_sets == rd["sepsets"]).all()
    for l, st in enumerate(dd.stats):
        assert st["executed"] == rd["executed"][l]


def test_admission_by_footprint_serialises_and_rejects():
    from gpucsl_amd.cli.cli_util import GpucslError
    from gpucsl_amd.serving import Request, SkeletonServer, request_footprint

    reqs = _requests(11, 4, 2)
    biggest = max(request_footprint(r) for r in reqs)
    srv = SkeletonServer(hbm_budget=biggest)      # room for about one job at a time
    for r in reqs:
        srv.submit(r)
    peak = 0
    done = []
    while not srv.idle:
        got = srv.tick()
        done += got
        peak = max(peak, srv.used)
        assert srv.used <= srv.budget
        if not got and not srv.launched:
            srv.wait()
    assert len(done) == len(reqs)
    for c in done:
        r = next(x for x in reqs if x.rid == c.rid)
        assert _same(c.result, _alone(r), r.kind)
    big = Request("gaussian", 100, 2000, 3, 0.01, X=torch.zeros((100, 2000), device="cuda"))
    with pytest.raises(GpucslError):
        srv.submit(big)


def test_served_float64_request_keeps_float64():
    """A Gaussian request with float64 data is served on the float64 path:
    its results equal the same request alone, and the FP64 oracle's skeleton
    and sepsets on those float64 values."""
    from gpucsl_amd.serving import Request, SkeletonServer

    X = syn.gaussian_sem(50, 2000, 3.0, seed=3).data
    assert X.dtype == np.float64
    req = Request("gaussian", 2000, 50, max_level=3, alpha=0.01, X=torch.from_numpy(X).cuda())
    srv = SkeletonServer(hbm_budget=1 << 30, host_budget=64 << 20, copy_results=True)
    srv.submit(req)
    out = srv.drain()
    assert len(out) == 1
    served = out[0].result
    alone = _alone(Request("gaussian", 2000, 50, max_level=3, alpha=0.01, X=torch.from_numpy(X).cuda()))
    assert _same(served, alone, "gaussian")
    ref = oracle.gauss_skeleton(oracle.corrcoef(X), 2000, 0.01, 3)
    assert (served.skeleton == ref["skeleton"]).all()
    assert (served.separation_sets == ref["sepsets"]).all()


def test_serving_stream_at_full_sizes():
    """The serving workload's request mix (32 requests of the seeded stream:
    70 % Gaussian p ~ U[20, 2000], n ~ U[500, 50000]; 30 % discrete p ~ U[20,
    300], n ~ U[1000, 100000]) plus a Gaussian p=2000, n=50000 and a discrete
    p=300, n=100000 request, all in flight together: every result is the same
    bits as the request alone; three Gaussian requests with p <= 310 and the
    discrete ones with p <= 110 equal the FP64 oracle."""
    import bench
    from gpucsl_amd.serving import Request, SkeletonServer

    reqs = bench.serving_requests(32, seed=500)
    reqs.append(Request("gaussian", 50000, 2000, 3, 0.01, X=syn.gaussian_sem_device(2000, 50000, 3.0, seed=777)))
    codes, dims = syn.discrete_bn_device(300, 100000, 2.0, seed=778)
    reqs.append(Request("discrete", 100000, 300, 3, 0.01, codes=codes, dims=dims))
    torch.cuda.synchronize()
    say(f"stream: {sum(r.kind == 'gaussian' for r in reqs)} Gaussian, {sum(r.kind == 'discrete' for r in reqs)} "
        f"discrete; largest Gaussian p={max(r.p for r in reqs if r.kind == 'gaussian')}")
    srv = SkeletonServer(copy_results=True)
    for r in reqs:
        srv.submit(r)
    done = srv.drain()
    assert len(done) == len(reqs) and srv.idle
    by_id = {c.rid: c for c in done}
    checked = 0
    for r in reqs:
        served = by_id[r.rid].result
        assert _same(served, _alone(r), r.kind), f"request {r.rid} ({r.kind} p={r.p} n={r.n}) differs"
        if r.kind == "gaussian" and r.p <= 310 and checked < 3:
            X = r.X.double().cpu().numpy()
            ref = oracle.gauss_skeleton(oracle.corrcoef(X), r.n, 0.01, r.max_level)
            assert (served.skeleton == ref["skeleton"]).all() and (served.separation_sets == ref["sepsets"]).all()
            checked += 1
        elif r.kind == "discrete" and r.p <= 110 and checked < 5:
            dims_h = r.dims.cpu().numpy().astype(np.uint8)
            col = np.ascontiguousarray(r.codes.cpu().numpy()[:, : r.n])
            ref = oracle.disc_skeleton(col, dims_h, 0.01, r.max_level)
            assert (served.skeleton == ref["skeleton"]).all() and (served.separation_sets == ref["sepsets"]).all()
            checked += 1
    say(f"stream: {len(reqs)} served equal alone; {checked} checked against the oracle; "
        f"{srv.ticks} ticks, largest batch {srv.max_batch}")
