"""gpucsl_gauss_skeleton_noalloc (workspace in the library's per-device array)
against gpucsl_gauss_skeleton with a caller workspace: the same bits, on two
calls in a row, the second with a smaller dataset; sizes beyond the array are
refused before anything runs.  GPU box only."""
import sys

import numpy as np
import pytest

from gpucsl_amd import _native as nat
from gpucsl_amd import synthetic as syn

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _buffers(p, L):
    dev = "cuda"
    return dict(G=torch.empty((p, p), dtype=torch.uint8, device=dev),
                sep=torch.empty((p, p, max(L, 1)), dtype=torch.int32, device=dev),
                zmin=torch.empty((p, p), dtype=torch.float32, device=dev),
                stats=torch.zeros(((L + 1) * nat.STAT_SLOTS,), dtype=torch.int64, device=dev))


def _run(entry, C32, C64, p, L, thr, ws=None):
    b = _buffers(p, L)
    args = [C32.data_ptr(), C64.data_ptr(), p, L, thr.ctypes.data, 8.0, 16, b["G"].data_ptr(),
            b["sep"].data_ptr(), b["zmin"].data_ptr(), b["stats"].data_ptr()]
    if ws is not None:
        args += [ws.data_ptr(), ws.numel()]
    nat.call(entry, *args, nat.stream_handle())
    return b


def _correlation(p, n, seed):
    from gpucsl_amd.gaussian import GaussianEngine

    X = syn.gaussian_sem(p, n, 4.0, seed=seed).data.astype(np.float32)
    eng = GaussianEngine(p, n, 1, [0.1, 0.1])
    eng.set_data(X)
    return eng.C32.clone(), eng.C64.clone()


@pytest.mark.parametrize("shapes", [[(1000, 10000, 4, 100), (300, 3000, 3, 7)]])
def test_noalloc_equals_workspace_entry_point(shapes):
    from gpucsl_amd.pc.helpers import get_gaussian_thresholds

    lib = nat.load()
    for call, (p, n, L, seed) in enumerate(shapes):
        say(f"[noalloc] call {call}: p={p} n={n} max_level={L} seed={seed}")
        C32, C64 = _correlation(p, n, seed)
        thr = np.ascontiguousarray(get_gaussian_thresholds(np.zeros((n, 1)), 0.01)[: L + 1], np.float64)
        ws = torch.empty((int(lib.gpucsl_gauss_workspace_bytes(p, L)),), dtype=torch.uint8, device="cuda")
        want = _run("gpucsl_gauss_skeleton", C32, C64, p, L, thr, ws)
        got = _run("gpucsl_gauss_skeleton_noalloc", C32, C64, p, L, thr)
        torch.cuda.synchronize()
        for k in ("G", "sep", "zmin", "stats"):
            assert torch.equal(got[k], want[k]), (call, k)
        edges = int(got["G"].sum().item()) // 2
        say(f"[noalloc] call {call}: identical, {edges} edges, "
            f"{int(got['stats'].view(-1, nat.STAT_SLOTS)[:, 6].sum().item())} tests")


def test_noalloc_rejects_sizes_beyond_its_array():
    lib = nat.load()
    p = 1235
    assert lib.gpucsl_gauss_workspace_bytes(p, 4) > nat.NOALLOC_SCRATCH_BYTES
    dummy = torch.zeros((1,), dtype=torch.float64, device="cuda")
    stats = torch.zeros((5 * nat.STAT_SLOTS,), dtype=torch.int64, device="cuda")
    thr = np.zeros(5, np.float64)
    rc = lib.gpucsl_gauss_skeleton_noalloc(dummy.data_ptr(), dummy.data_ptr(), p, 4, thr.ctypes.data, 8.0, 16,
                                           dummy.data_ptr(), dummy.data_ptr(), dummy.data_ptr(), stats.data_ptr(),
                                           nat.stream_handle())
    assert rc == 4
    assert b"scratch" in lib.gpucsl_last_error()
