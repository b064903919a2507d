"""Device-resident discrete PC-stable skeleton discovery for one dataset.

DiscreteEngine owns the HBM buffers of one job (padded column-major uint8
codes, category counts, adjacency, separation sets, pmax, statistics and the
workspace) and drives libgpucsl_hip.so on the current torch stream.  This is
the device part of gpucsl/pc/discover_skeleton_discrete.py:24-128.
"""
from __future__ import annotations

import ctypes as ct
from dataclasses import dataclass

import numpy as np

from gpucsl_amd import _native as nat
from gpucsl_amd.cli.cli_util import GpucslError

COLUMN_ALIGN = 64   # bytes; a multiple of 16 as the kernels require
BATCH_JOBS = 16     # GPUCSL_DISC_BATCH_JOBS: jobs per batched launch


class DiscJob(ct.Structure):
    """ctypes mirror of struct gpucsl_disc_job (include/gpucsl_hip.h)."""
    _fields_ = [
        ("data", ct.c_void_p), ("dims", ct.c_void_p), ("n", ct.c_int64), ("stride", ct.c_int64),
        ("p", ct.c_int32), ("max_level", ct.c_int32), ("max_dim", ct.c_int32), ("reserved", ct.c_int32),
        ("alpha", ct.c_double), ("G", ct.c_void_p), ("sepsets", ct.c_void_p), ("pmax", ct.c_void_p),
        ("stats", ct.c_void_p), ("skeleton", ct.c_void_p), ("ws", ct.c_void_p), ("ws_bytes", ct.c_size_t),
    ]


def ordinal_encode(data: np.ndarray):
    """Per-column codes 0..k-1 in sorted order of the distinct values, as
    sklearn's OrdinalEncoder (gpucsl/pc/discover_skeleton_discrete.py:38-39).
    Returns (codes uint8 (n, p), dims uint8 (p,))."""
    data = np.asarray(data)
    n, p = data.shape
    codes = np.empty((n, p), np.uint8)
    dims = np.empty(p, np.int64)
    for v in range(p):
        uniq, inv = np.unique(data[:, v], return_inverse=True)
        if uniq.size > 255:
            raise GpucslError(f"variable {v} has {uniq.size} categories; at most 255 are supported")
        codes[:, v] = inv.reshape(-1)
        dims[v] = uniq.size
    return codes, dims.astype(np.uint8)


# numpy dtype -> (dtype code of gpucsl_disc_encode, numpy dtype it is passed as)
def _encode_dtype(a: np.ndarray):
    k, size = a.dtype.kind, a.dtype.itemsize
    if k == "b" or (k in "iu" and size <= 2) or (k == "i" and size == 4):
        return 0, np.int32
    if (k == "i" and size == 8) or (k == "u" and size == 4):
        return 1, np.int64
    if k == "u" and size == 8 and (a.size == 0 or int(a.max()) < (1 << 63)):
        return 1, np.int64
    if k == "f" and size <= 4:
        return 2, np.float32
    if k == "f" and size == 8:
        return 3, np.float64
    return None


def ordinal_encode_device(data, device=None):
    """ordinal_encode on the GPU for numeric data: the values go to the device
    as they are and gpucsl_disc_encode codes them there.  Returns (codes (p,
    stride) uint8 tensor, padded column-major; dims (p,) uint8 tensor; dims as a
    host array), or None for data it does not take (non-numeric, or NaN, which
    the host encoder treats as a category)."""
    torch = nat.require_gpu()
    a = np.asarray(data)
    if a.ndim != 2 or a.size == 0:
        return None
    kind = _encode_dtype(a)
    if kind is None:
        return None
    code, np_dtype = kind
    n, p = a.shape
    dev = device if device is not None else torch.device("cuda", torch.cuda.current_device())
    stride = (n + COLUMN_ALIGN - 1) // COLUMN_ALIGN * COLUMN_ALIGN
    vals = torch.from_numpy(np.ascontiguousarray(a, dtype=np_dtype)).to(dev)
    codes = torch.empty((p, stride), dtype=torch.uint8, device=dev)
    dims = torch.empty((p,), dtype=torch.uint8, device=dev)
    status = torch.empty((p,), dtype=torch.int32, device=dev)
    nat.call("gpucsl_disc_encode", vals.data_ptr(), code, n, p, p, 1, codes.data_ptr(), stride, dims.data_ptr(),
             status.data_ptr(), nat.stream_handle())
    st = status.cpu().numpy()
    if (st == 1).any():
        v = int(np.flatnonzero(st == 1)[0])
        raise GpucslError(f"variable {v} has more than 255 categories; at most 255 are supported")
    if (st == 2).any():
        return None
    return codes, dims, dims.cpu().numpy()


@dataclass
class DiscreteResult:
    skeleton: np.ndarray          # (p, p) uint16
    separation_sets: np.ndarray   # (p, p, max_level) int32
    pmax: np.ndarray              # (p, p) float32, -1 on and below the diagonal
    stats: list


def disc_workspace(p: int, max_level: int, max_dim: int, limit: int | None = None) -> int:
    """Workspace bytes: the full size (most tests at once), or `limit` when it
    is smaller but still holds the minimum.  RuntimeError when the tables of
    max_level can never fit (gpucsl/pc/kernel_management.py:434-442)."""
    lib = nat.load()
    full = int(lib.gpucsl_disc_workspace_bytes(p, max_level, max(max_dim, 1)))
    least = int(lib.gpucsl_disc_min_workspace_bytes(p, max_level, max(max_dim, 1)))
    if least == (1 << 64) - 1 or (limit is not None and limit < least):
        raise RuntimeError(f"Out of memory for maximum level of {max_level}: the workspace needs at least "
                           f"{least} bytes" + (f", the restriction is {limit}" if limit is not None else ""))
    return full if limit is None or limit >= full else int(limit)


def disc_layout(n: int, p: int, max_level: int, max_dim: int, ws_limit: int | None = None):
    """Byte offsets of one discrete job's buffers in a single arena, and its size."""
    stride = (n + COLUMN_ALIGN - 1) // COLUMN_ALIGN * COLUMN_ALIGN
    L = max(max_level, 1)
    pp = p * p
    sizes = [("data", p * stride), ("dims", p), ("G", pp), ("sep", 4 * pp * L), ("pmax", 4 * pp),
             ("skel16", 2 * pp), ("stats", 8 * (max_level + 1) * nat.STAT_SLOTS),
             ("ws", disc_workspace(p, max_level, max_dim, ws_limit))]
    off, out = 0, {}
    for name, sz in sizes:
        out[name] = (off, sz)
        off = (off + sz + 255) // 256 * 256
    return out, off


class DiscreteEngine:
    def __init__(self, p: int, n: int, max_level: int, alpha: float, max_dim: int, arena=None,
                 ws_limit: int | None = None):
        """arena: optional uint8 device tensor of at least disc_layout(...)[1]
        bytes that holds every buffer of the job (the serving pool).
        ws_limit: cap on the workspace bytes (fewer tests run at once)."""
        torch = nat.require_gpu()
        if p < 2:
            raise GpucslError("need at least two variables")
        self.torch = torch
        self.p, self.n, self.max_level, self.alpha = int(p), int(n), int(max_level), float(alpha)
        self.max_dim = int(max(max_dim, 1))
        dev = torch.device("cuda", torch.cuda.current_device())
        self.device = dev
        self.stride = (self.n + COLUMN_ALIGN - 1) // COLUMN_ALIGN * COLUMN_ALIGN
        L = max(self.max_level, 1)
        layout, nbytes = disc_layout(self.n, self.p, self.max_level, self.max_dim, ws_limit)
        if arena is None:
            arena = torch.empty((nbytes,), dtype=torch.uint8, device=dev)
        elif arena.numel() < nbytes:
            raise GpucslError("arena too small for the discrete job")
        self.arena, self.nbytes = arena, nbytes

        def view(name, dtype, shape):
            off, sz = layout[name]
            return arena[off:off + sz].view(dtype).view(*shape)

        self.data = view("data", torch.uint8, (p, self.stride))
        self.data.zero_()
        self.dims = view("dims", torch.uint8, (p,))
        self.G = view("G", torch.uint8, (p, p))
        self.sep = view("sep", torch.int32, (p, p, L))
        self.pmax = view("pmax", torch.float32, (p, p))
        self.skel16 = view("skel16", torch.int16, (p, p))
        self.stats = view("stats", torch.int64, ((self.max_level + 1) * nat.STAT_SLOTS,))
        self.stats.zero_()
        self.ws_bytes = layout["ws"][1]
        self.ws = view("ws", torch.uint8, (self.ws_bytes,))

    def workspace_bytes(self) -> int:
        return self.ws_bytes

    def set_codes(self, codes: np.ndarray, dims: np.ndarray) -> None:
        """codes: (n, p) uint8 ordinal codes; dims: (p,) category counts."""
        torch = self.torch
        codes = np.asarray(codes, np.uint8)
        if codes.shape != (self.n, self.p):
            raise GpucslError(f"codes must be {self.n}x{self.p}")
        if int(np.max(dims)) > self.max_dim:
            raise GpucslError("a variable has more categories than max_dim")
        host = np.zeros((self.p, self.stride), np.uint8)
        host[:, : self.n] = codes.T
        self.data.copy_(torch.from_numpy(host))
        self.dims.copy_(torch.from_numpy(np.ascontiguousarray(dims, np.uint8)))

    def set_codes_device(self, colmajor, dims) -> None:
        """Device input: colmajor (p, >= n) uint8 codes, one row per variable
        (only the first n columns are read), dims (p,) category counts."""
        torch = self.torch
        if colmajor.dtype != torch.uint8 or colmajor.shape[0] != self.p or colmajor.shape[1] < self.n:
            raise GpucslError(f"codes must be uint8 ({self.p}, >= {self.n})")
        self.data[:, : self.n].copy_(colmajor[:, : self.n], non_blocking=True)
        self.dims.copy_(dims.to(torch.uint8), non_blocking=True)

    # ---- level by level (the serving loop interleaves these with other jobs)
    def start(self) -> None:
        """Buffer set-up and level 0."""
        s = nat.stream_handle()
        nat.call("gpucsl_disc_init", self.p, self.max_level, self.max_dim, self.G.data_ptr(), self.sep.data_ptr(),
                 self.pmax.data_ptr(), self.stats.data_ptr(), self.ws.data_ptr(), self.ws_bytes, s)
        nat.call("gpucsl_disc_level0", self.data.data_ptr(), self.n, self.stride, self.dims.data_ptr(), self.p,
                 self.max_level, self.max_dim, self.alpha, self.G.data_ptr(), self.pmax.data_ptr(),
                 self.stats.data_ptr(), self.ws.data_ptr(), self.ws_bytes, s)

    def level(self, l: int) -> None:
        st = self.stats.data_ptr() + 8 * nat.STAT_SLOTS * l
        nat.call("gpucsl_disc_level", l, self.data.data_ptr(), self.n, self.stride, self.dims.data_ptr(), self.p,
                 self.max_level, self.max_dim, self.alpha, self.G.data_ptr(), self.sep.data_ptr(),
                 self.pmax.data_ptr(), st, self.ws.data_ptr(), self.ws_bytes, nat.stream_handle())

    def finish(self) -> None:
        nat.call("gpucsl_disc_finish", self.G.data_ptr(), self.p, self.pmax.data_ptr(), self.skel16.data_ptr(),
                 nat.stream_handle())

    def run(self) -> None:
        nat.call("gpucsl_disc_skeleton", self.data.data_ptr(), self.n, self.stride, self.dims.data_ptr(),
                 self.p, self.max_level, self.max_dim, self.alpha, self.G.data_ptr(), self.sep.data_ptr(),
                 self.pmax.data_ptr(), self.stats.data_ptr(), self.ws.data_ptr(), self.ws_bytes,
                 nat.stream_handle())
        nat.call("gpucsl_skeleton_to_u16", self.G.data_ptr(), self.p, self.skel16.data_ptr(),
                 nat.stream_handle())

    def level_stats(self):
        return nat.decode_stats(self.stats.cpu().numpy().view(np.uint64), self.max_level)

    def descriptor(self) -> DiscJob:
        """This job's entry in a batched launch (its own buffers and workspace)."""
        return DiscJob(self.data.data_ptr(), self.dims.data_ptr(), self.n, self.stride, self.p, self.max_level,
                       self.max_dim, 0, self.alpha, self.G.data_ptr(), self.sep.data_ptr(), self.pmax.data_ptr(),
                       self.stats.data_ptr(), self.skel16.data_ptr(), self.ws.data_ptr(), self.ws_bytes)

    def results(self) -> DiscreteResult:
        skel, sep, pmax, st = nat.to_host(self.skel16, self.sep, self.pmax, self.stats)
        return DiscreteResult(skel.view(np.uint16), sep[:, :, : self.max_level], pmax,
                              nat.decode_stats(st.view(np.uint64), self.max_level))

    def ci_test(self, vi: int, vj: int, S) -> tuple:
        """One CI test on the loaded data: (table, chi2, p, df)."""
        torch = self.torch
        S = list(S)
        l = len(S)
        dS = torch.tensor(S if S else [0], dtype=torch.int32, device=self.device)
        dims = self.dims.cpu().numpy()
        cells = int(dims[vi]) * int(dims[vj]) * int(np.prod([int(dims[s]) for s in S], dtype=np.int64))
        table = torch.zeros((max(cells, 1),), dtype=torch.int32, device=self.device)
        out = torch.zeros((3,), dtype=torch.float64, device=self.device)
        nat.call("gpucsl_disc_ci_test", self.data.data_ptr(), self.n, self.stride, self.dims.data_ptr(), vi, vj,
                 dS.data_ptr(), l, 1 if l == 0 else 0, table.data_ptr(), out.data_ptr(), nat.stream_handle())
        o = out.cpu().numpy()
        if o[2] < 0:
            raise GpucslError(f"ci_test: the table of {cells} cells and its marginals exceed one workgroup's LDS")
        return table.cpu().numpy().view(np.uint32)[:cells].copy(), float(o[0]), float(o[1]), int(o[2])


class DiscreteBatch:
    """Discrete jobs (DiscreteEngine, each with its own buffers) advancing
    together through the batched entry points: one launch per kernel of a
    level for up to BATCH_JOBS jobs, where the single-dataset path makes one per
    job.  No counterpart in gpucsl, which runs
    gpucsl/pc/discover_skeleton_discrete.py:24-128 once per dataset; each job's
    results are bit-identical to its single-dataset run.  The whole skeleton
    can be captured into one HIP graph (``capture``) and replayed."""

    def __init__(self, engines: list):
        torch = nat.require_gpu()
        if not engines:
            raise GpucslError("empty batch")
        self.torch = torch
        self.engines = list(engines)
        self.max_level = max(e.max_level for e in self.engines)
        self.table = (DiscJob * len(self.engines))(*[e.descriptor() for e in self.engines])
        self._graph = None

    def _args(self):
        return ct.byref(self.table), len(self.engines)

    # ---- level by level
    def start(self) -> None:
        """Buffer set-up and level 0 of every job."""
        s = nat.stream_handle()
        nat.call("gpucsl_disc_batch_init", *self._args(), s)
        nat.call("gpucsl_disc_batch_level0", *self._args(), s)

    def level(self, l: int) -> None:
        """Level l of every job whose max_level reaches it."""
        nat.call("gpucsl_disc_batch_level", l, *self._args(), nat.stream_handle())

    def finish(self) -> None:
        """pmax lower triangle and the uint16 skeletons."""
        nat.call("gpucsl_disc_batch_finish", *self._args(), nat.stream_handle())

    def launch(self) -> None:
        """Set-up, levels 0..max_level and the finish, no host sync."""
        nat.call("gpucsl_disc_batch_skeleton", *self._args(), nat.stream_handle())

    def capture(self) -> None:
        """Capture ``launch`` into one HIP graph."""
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

    def results(self) -> list:
        return [e.results() for e in self.engines]
