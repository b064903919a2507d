"""Many Gaussian skeleton jobs (datasets or bootstrap resamples of uneven p and
n) advancing together, one shared launch per kernel and level.

The reference runs gaussian_ci_worker once per dataset
(gpucsl/pc/discover_skeleton_gaussian.py:30-113); a caller with many datasets
loops over it, paying every launch and host synchronisation per dataset.  Here
each job keeps its HBM state in a ``GaussJobState`` (X, C32/C64, adjacency,
separation sets, zmin, pmax and the level workspace), and a ``GaussianBatch``
hands a table of job descriptors (``gpucsl_gauss_job`` in include/gpucsl_hip.h)
to the batched entry points: the level kernels walk the concatenated work
items of all jobs, so small jobs fill the GPU together.

A batch can be captured once into a HIP graph (``capture``) and replayed; the
whole skeleton (correlation, levels 0..L, pmax) is then one graph launch.
The results of a job do not depend on which other jobs share its batch.
"""
from __future__ import annotations

import ctypes as ct
from dataclasses import dataclass

import numpy as np

from gpucsl_amd import _native as nat
from gpucsl_amd.cli.cli_util import GpucslError
from gpucsl_amd.gaussian import DEFAULT_GROUP_SIZE, DEFAULT_MARGIN, GaussianResult, thresholds_for_levels

MAX_LEVELS = 33         # GPUCSL_MAX_LEVELS: thresholds for levels 0..32 (GPUCSL_MAX_LEVEL)
ALIGN = 256


class GaussJob(ct.Structure):
    """ctypes mirror of struct gpucsl_gauss_job (include/gpucsl_hip.h)."""
    _fields_ = [
        ("p", ct.c_int32), ("max_level", ct.c_int32), ("n", ct.c_int64),
        ("x_dtype", ct.c_int32), ("k_splits", ct.c_int32), ("flags", ct.c_int32), ("reserved", ct.c_int32),
        ("X", ct.c_void_p), ("C32", ct.c_void_p), ("C64", ct.c_void_p), ("G", ct.c_void_p),
        ("sepsets", ct.c_void_p), ("zmin", ct.c_void_p), ("pmax", ct.c_void_p),
        ("skeleton", ct.c_void_p), ("ws", ct.c_void_p), ("corr_ws", ct.c_void_p),
        ("thr", ct.c_double * MAX_LEVELS),
    ]


PCALG_COMPAT = 1   # GPUCSL_JOB_PCALG_COMPAT


def pcalg_thresholds(n: int, alpha: float, levels: int = MAX_LEVELS) -> np.ndarray:
    """pcalg's |Phi^-1(alpha/2)| / sqrt(n - l - 3), 0.0 where n - l - 3 <= 0."""
    from scipy.stats import norm

    q = abs(norm.ppf(alpha / 2))
    return np.array([q / np.sqrt(n - l - 3) if n - l - 3 > 0 else 0.0 for l in range(levels)], np.float64)


def fisher_z_thresholds(n: int, alpha: float, levels: int = MAX_LEVELS) -> np.ndarray:
    """|Phi^-1(alpha/2)| / sqrt(n - l - 4) for l = 0..levels-1, 0.0 where the
    reference leaves the entry unset (gpucsl/pc/helpers.py:42-50)."""
    from scipy.stats import norm

    q = abs(norm.ppf(alpha / 2))
    thr = np.zeros(levels, np.float64)
    for i in range(1, min(50, n - 3) - 1):
        if i - 1 < levels:
            thr[i - 1] = q / np.sqrt(n - i - 3)
    return thr


def _al(x: int) -> int:
    return (x + ALIGN - 1) // ALIGN * ALIGN


def job_layout(n: int, p: int, max_level: int, x_itemsize: int = 4, k_splits: int | None = None):
    """Byte offsets of one job's buffers in its arena and the arena size.

    Footprint = n p x_itemsize (X) + p^2 (4 C32 + 8 C64 + 1 G + 4 zmin + 4 pmax
    + 2 skeleton) + 4 p^2 L (sepsets) + level and correlation workspaces."""
    lib = nat.load()
    if k_splits is None:
        k_splits = int(lib.gpucsl_corr_k_splits(n, p, 0 if x_itemsize == 4 else 1))
    L = max(max_level, 1)
    pp = p * p
    sizes = [
        ("X", n * p * x_itemsize),
        ("C32", 4 * pp), ("C64", 8 * pp), ("G", pp), ("sepsets", 4 * pp * L),
        ("zmin", 4 * pp), ("pmax", 4 * pp), ("skeleton", 2 * pp),
        ("ws", int(lib.gpucsl_gauss_job_workspace_bytes(p))),
        ("corr_ws", int(lib.gpucsl_corr_job_workspace_bytes(n, p, k_splits))),
    ]
    off, out = 0, {}
    for name, sz in sizes:
        out[name] = (off, sz)
        off = _al(off + sz)
    return out, off


def job_footprint(n: int, p: int, max_level: int, x_itemsize: int = 4, k_splits: int | None = None) -> int:
    return job_layout(n, p, max_level, x_itemsize, k_splits)[1]


class GaussJobState:
    """HBM state of one Gaussian job inside one arena (a uint8 torch tensor)."""

    def __init__(self, n: int, p: int, max_level: int, alpha: float | None = None, thresholds=None,
                 x_dtype: str = "float32", k_splits: int | None = None, device=None, arena=None,
                 pcalg_compat: bool = False):
        torch = nat.require_gpu()
        if p < 2 or n < 5:
            raise GpucslError(f"job needs p >= 2 and n >= 5 (got p={p}, n={n})")
        if max_level < 0:
            raise GpucslError("max_level must be >= 0")
        if x_dtype not in ("float32", "float64"):
            raise GpucslError("x_dtype must be float32 or float64")
        self.torch = torch
        self.n, self.p, self.max_level = int(n), int(p), int(max_level)
        self.x_dtype = x_dtype
        # split-K of the correlation GEMM: a function of (n, p) only, so that the
        # job's correlation matrix (and everything after it) is the same bits
        # whatever batch it runs in, and the same as the single-dataset path's
        if k_splits is None:
            k_splits = int(nat.load().gpucsl_corr_k_splits(self.n, self.p, 0 if x_dtype == "float32" else 1))
        self.k_splits = max(1, int(k_splits))
        # pcalg conventions: thresholds sqrt(n - l - 3), pMax over every test,
        # one-way sepsets (GPUCSL_JOB_PCALG_COMPAT)
        self.flags = PCALG_COMPAT if pcalg_compat else 0
        if thresholds is None:
            if alpha is None:
                raise GpucslError("give alpha or thresholds")
            thresholds = (pcalg_thresholds if pcalg_compat else fisher_z_thresholds)(self.n, float(alpha))
        self.thr = thresholds_for_levels(thresholds, self.max_level)
        isz = 4 if x_dtype == "float32" else 8
        self.layout, self.nbytes = job_layout(self.n, self.p, self.max_level, isz, self.k_splits)
        dev = device if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.device = dev
        if arena is None:
            arena = torch.empty((self.nbytes,), dtype=torch.uint8, device=dev)
        elif arena.numel() < self.nbytes:
            raise GpucslError("arena too small for the job")
        self.arena = arena
        self.has_x = False
        self.has_c64 = False

    def view(self, name, dtype, shape):
        off, sz = self.layout[name]
        t = self.arena[off:off + sz].view(dtype)
        return t.view(*shape)

    def ptr(self, name) -> int:
        return self.arena.data_ptr() + self.layout[name][0]

    # ----------------------------------------------------------------- input
    def set_data(self, X) -> None:
        """Copy X (n x p) into the job's HBM (stream-ordered on the current stream)."""
        torch = self.torch
        want = torch.float32 if self.x_dtype == "float32" else torch.float64
        if not isinstance(X, torch.Tensor):
            X = torch.from_numpy(np.ascontiguousarray(X))
        if tuple(X.shape) != (self.n, self.p):
            raise GpucslError(f"X must be {self.n}x{self.p}, got {tuple(X.shape)}")
        self.view("X", want, (self.n, self.p)).copy_(X.to(want), non_blocking=True)
        self.has_x = True

    def set_correlation(self, C) -> None:
        torch = self.torch
        if not isinstance(C, torch.Tensor):
            C = torch.from_numpy(np.ascontiguousarray(np.asarray(C, np.float64)))
        if tuple(C.shape) != (self.p, self.p):
            raise GpucslError(f"correlation matrix must be {self.p}x{self.p}")
        self.view("C64", torch.float64, (self.p, self.p)).copy_(C.to(torch.float64), non_blocking=True)
        self.has_c64 = True
        self.has_x = False

    # ------------------------------------------------------------ descriptor
    def descriptor(self) -> GaussJob:
        j = GaussJob()
        j.p, j.max_level, j.n = self.p, self.max_level, self.n
        j.x_dtype = 0 if self.x_dtype == "float32" else 1
        j.k_splits = self.k_splits
        j.flags = self.flags
        j.X = self.ptr("X") if self.has_x else None
        for name in ("C32", "C64", "G", "sepsets", "zmin", "pmax", "skeleton", "ws", "corr_ws"):
            setattr(j, name, self.ptr(name))
        for l in range(MAX_LEVELS):
            j.thr[l] = float(self.thr[l]) if l < self.thr.size else 0.0
        return j

    # ---------------------------------------------------------------- output
    def results(self, stats=None) -> GaussianResult:
        torch = self.torch
        p, L = self.p, max(self.max_level, 1)
        skel, sep, pmax, zmin = nat.to_host(self.view("skeleton", torch.int16, (p, p)),
                                            self.view("sepsets", torch.int32, (p, p, L)),
                                            self.view("pmax", torch.float32, (p, p)),
                                            self.view("zmin", torch.float32, (p, p)))
        return GaussianResult(skel.view(np.uint16), sep[:, :, : self.max_level], pmax, zmin,
                              stats if stats is not None else [])


@dataclass
class BatchStats:
    levels: list   # per-level dicts summed over the batch (executed, total, useful, ...)

    @property
    def useful(self) -> int:
        return sum(d["useful"] for d in self.levels)

    @property
    def executed(self) -> int:
        return sum(d["executed"] for d in self.levels)


class GaussianBatch:
    """A fixed set of jobs driven through the batched entry points."""

    def __init__(self, jobs: list, margin: float = DEFAULT_MARGIN, group_size: int = DEFAULT_GROUP_SIZE):
        torch = nat.require_gpu()
        if not jobs:
            raise GpucslError("empty batch")
        self.torch = torch
        self.jobs = list(jobs)
        self.margin = float(margin)
        self.group_size = int(group_size)
        self.max_level = max(j.max_level for j in self.jobs)
        dev = self.jobs[0].device
        self.device = dev
        self.stats = torch.zeros(((self.max_level + 1) * nat.STAT_SLOTS,), dtype=torch.int64, device=dev)
        self._graph = None
        self.refresh()

    def refresh(self) -> None:
        """Rebuild the job table (host and device) from the jobs' current state."""
        torch = self.torch
        nj = len(self.jobs)
        self.table = (GaussJob * nj)(*[j.descriptor() for j in self.jobs])
        raw = np.frombuffer(bytes(self.table), np.uint8)
        self.table_dev = torch.from_numpy(raw.copy()).to(self.device)
        lib = nat.load()
        self.ws_bytes = int(lib.gpucsl_gauss_batch_workspace_bytes(ct.byref(self.table), nj))
        self.ws = torch.empty((self.ws_bytes,), dtype=torch.uint8, device=self.device)
        self._graph = None

    def _args(self):
        return ct.byref(self.table), self.table_dev.data_ptr(), len(self.jobs)

    def launch(self) -> None:
        """corr + init + all levels + pmax/skeleton, stream-ordered, no host sync."""
        for j in self.jobs:
            if not (j.has_x or j.has_c64):
                raise GpucslError("every job needs data or a correlation matrix before run")
        xs = {j.has_x for j in self.jobs}
        if len(xs) != 1:
            raise GpucslError("a batch mixes jobs with data and jobs with a correlation matrix")
        nat.call("gpucsl_gauss_batch_skeleton", *self._args(), self.margin, self.group_size,
                 self.stats.data_ptr(), self.ws.data_ptr(), self.ws_bytes, nat.stream_handle())

    def capture(self) -> None:
        """Capture ``launch`` into one HIP graph (the whole skeleton)."""
        torch = self.torch
        self.launch()                     # warm the code objects outside capture
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            g.capture_begin()
            self.launch()
            g.capture_end()
        torch.cuda.current_stream().wait_stream(s)
        self._graph = g

    def run(self) -> None:
        if self._graph is not None:
            self._graph.replay()
        else:
            self.launch()

    def batch_stats(self) -> BatchStats:
        host = self.stats.cpu().numpy().view(np.uint64)
        return BatchStats(nat.decode_stats(host, self.max_level))

    def results(self) -> list:
        return [j.results() for j in self.jobs]


def run_plain(jobs: list, margin: float = DEFAULT_MARGIN, group_size: int = DEFAULT_GROUP_SIZE) -> None:
    """The same jobs one after another through the single-dataset entry points
    (the drop-in path), on the current stream: the comparison the batched path
    is measured against."""
    torch = nat.require_gpu()
    lib = nat.load()
    s = nat.stream_handle()
    for j in jobs:
        p = j.p
        ws_bytes = int(lib.gpucsl_gauss_workspace_bytes(p, j.max_level))
        if not hasattr(j, "_plain_ws") or j._plain_ws.numel() < ws_bytes:
            j._plain_ws = torch.empty((ws_bytes,), dtype=torch.uint8, device=j.device)
            j._plain_stats = torch.zeros(((j.max_level + 1) * nat.STAT_SLOTS,), dtype=torch.int64,
                                         device=j.device)
        if j.has_x:
            cb = int(lib.gpucsl_corr_workspace_bytes(j.n, p))
            if not hasattr(j, "_plain_corr_ws") or j._plain_corr_ws.numel() < cb:
                j._plain_corr_ws = torch.empty((cb,), dtype=torch.uint8, device=j.device)
            nat.call("gpucsl_corr", j.ptr("X"), 0 if j.x_dtype == "float32" else 1, j.n, p, j.ptr("C64"),
                     j.ptr("C32"), j._plain_corr_ws.data_ptr(), cb, s)
        else:
            nat.call("gpucsl_corr_c64_to_c32", j.ptr("C64"), p, j.ptr("C32"), s)
        thr = np.ascontiguousarray(j.thr)
        nat.call("gpucsl_gauss_skeleton", j.ptr("C32"), j.ptr("C64"), p, j.max_level, thr.ctypes.data,
                 float(margin), int(group_size), j.ptr("G"), j.ptr("sepsets"), j.ptr("zmin"),
                 j._plain_stats.data_ptr(), j._plain_ws.data_ptr(), ws_bytes, s)
        nat.call("gpucsl_gauss_outputs", j.ptr("zmin"), j.ptr("G"), p, j.n, j.ptr("pmax"), j.ptr("skeleton"), s)
