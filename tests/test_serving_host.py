"""Host logic of the serving loop (CPU): the first-fit pool that admission
draws from, and the footprints it charges."""
import numpy as np
import pytest

from gpucsl_amd.serving import Pool, Request, request_footprint, result_bytes


def test_pool_first_fit_no_overlap_and_coalesces():
    rng = np.random.default_rng(0)
    pool = Pool(1 << 20)
    live = []
    for step in range(2000):
        if live and (rng.random() < 0.45 or len(live) > 40):
            k = int(rng.integers(len(live)))
            pool.release(*live.pop(k))
        else:
            r = pool.alloc(int(rng.integers(1, 60000)))
            if r is not None:
                live.append(r)
        spans = sorted(live)
        for (a, la), (b, _) in zip(spans, spans[1:]):
            assert a + la <= b                       # no overlap
        assert all(off % 256 == 0 and ln % 256 == 0 for off, ln in live)
        assert pool.in_use == sum(ln for _, ln in live)
        free = pool.free
        assert all(free[i][0] + free[i][1] < free[i + 1][0] for i in range(len(free) - 1))  # coalesced
    for r in live:
        pool.release(*r)
    assert pool.free == [(0, 1 << 20)] and pool.in_use == 0


def test_pool_refuses_what_does_not_fit():
    pool = Pool(4096)
    a = pool.alloc(3000)
    assert a == (0, 3072)
    assert pool.alloc(2000) is None
    assert pool.alloc(1024) == (3072, 1024)
    pool.release(*a)
    assert pool.alloc(3072) == (0, 3072)


def test_footprints_cover_the_device_state():
    g = Request("gaussian", 10000, 1000, 4)
    # X (n p 4) + C32 + C64 + G + sepsets + zmin + pmax + skeleton at least
    floor = 10000 * 1000 * 4 + 1000 * 1000 * (4 + 8 + 1 + 4 * 4 + 4 + 4 + 2)
    assert request_footprint(g) >= floor
    d = Request("discrete", 50000, 200, 3, max_dim=5)
    assert request_footprint(d) >= 200 * 50000 + 200 * 200 * (1 + 12 + 4 + 2)
    assert result_bytes(g) >= 1000 * 1000 * (2 + 16 + 4 + 4)
    with pytest.raises(Exception):
        request_footprint(Request("other", 10, 10))
