"""The kernels' host-callable bodies built for the host under ASan/UBSan
(tests/native/host_kernels.cc): workspace layouts stay inside their sizes with
disjoint regions, and the run-time unranking, the FP64 Fisher-z in the
reference's operation order and Perlman's pchisq equal the oracle bit for bit.
CPU only; the build takes about 20 s."""
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent
SRC = HERE / "native" / "host_kernels.cc"
DEPS = [SRC, ROOT / "gpucsl_amd" / "csrc" / "common.h", ROOT / "gpucsl_amd" / "csrc" / "job.h",
        ROOT / "include" / "gpucsl_hip.h"]


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(hipcc).exists():
        pytest.skip("hipcc not available")
    out = tmp_path_factory.mktemp("host_kernels") / "host_kernels"
    cmd = [hipcc, "-x", "hip", "--offload-host-only", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer",
           "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined", "-fno-sanitize-recover=all",
           str(SRC), "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    return out


def _run(exe, *args):
    r = subprocess.run([str(exe), *map(str, args)], capture_output=True, text=True,
                       env={"ASAN_OPTIONS": "detect_leaks=0:abort_on_error=0", "PATH": "/usr/bin:/bin"})
    assert r.returncode == 0, (r.returncode, r.stderr[-3000:])
    return r.stdout


def test_workspace_layouts(exe):
    assert "layout ok" in _run(exe, "layout")


@pytest.mark.parametrize("N,l", [(9, 9), (12, 9), (14, 10), (20, 12), (13, 3)])
def test_runtime_unrank_equals_oracle(exe, N, l):
    count = min(oracle.binomial(N, l), 3000)
    got = [list(map(int, ln.split())) for ln in _run(exe, "unrank", N, l, count).splitlines()]
    assert len(got) == count
    for r in range(count):
        assert got[r] == oracle.unrank(N, l, r).tolist(), r


def test_fp64_fisher_z_equals_oracle(exe, tmp_path):
    rng = np.random.default_rng(3)
    p = 20
    A = rng.standard_normal((200, p))
    C = np.corrcoef(A, rowvar=False)
    cases = []
    for _ in range(400):
        l = int(rng.integers(1, 9))
        idx = rng.choice(p, size=l + 2, replace=False)
        cases.append((int(idx[0]), int(idx[1]), l, sorted(int(x) for x in idx[2:])))
    with open(tmp_path / "in", "wb") as f:
        np.array([p, len(cases)], np.int32).tofile(f)
        C.astype(np.float64).tofile(f)
        for vi, vj, l, S in cases:
            np.array([vi, vj, l] + S + [0] * (8 - l), np.int32).tofile(f)
    _run(exe, "gauss_z", tmp_path / "in", tmp_path / "out")
    z = np.fromfile(tmp_path / "out", np.float64)
    want = np.array([oracle.gauss_z(C, vi, vj, S) for vi, vj, _, S in cases])
    assert np.array_equal(z.view(np.uint64), want.view(np.uint64))


def test_pchisq_equals_oracle(exe, tmp_path):
    rng = np.random.default_rng(4)
    xs = np.concatenate([rng.uniform(0, 60, 300), rng.uniform(0, 3000, 100), [0.0, 1e-9, 41.0, 2500.0]])
    dfs = np.concatenate([rng.integers(1, 40, 300), rng.integers(100, 2100, 100), [1, 2, 40, 2000]]).astype(np.int32)
    with open(tmp_path / "in", "wb") as f:
        np.array([xs.size], np.int32).tofile(f)
        for x, d in zip(xs, dfs):
            np.array([x], np.float64).tofile(f)
            np.array([d], np.int32).tofile(f)
    _run(exe, "pchisq", tmp_path / "in", tmp_path / "out")
    got = np.fromfile(tmp_path / "out", np.float64)
    want = np.array([oracle.pchisq(float(x), int(d)) for x, d in zip(xs, dfs)])
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))
