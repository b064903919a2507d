"""Device-resident discrete PC-stable skeleton discovery for one dataset.

DiscreteEngine owns the HBM buffers of one job (padded column-major uint8
codes, category counts, adjacency, separation sets, pmax, statistics and the
workspace) and drives libgpucsl_hip.so on the current torch stream.  This is
the device part of gpucsl/pc/discover_skeleton_discrete.py:24-128.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from gpucsl_amd import _native as nat
from gpucsl_amd.cli.cli_util import GpucslError

COLUMN_ALIGN = 64   # bytes; a multiple of 16 as the kernels require


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


@dataclass
class DiscreteResult:
    skeleton: np.ndarray          # (p, p) uint16
    separation_sets: np.ndarray   # (p, p, max_level) int32
    pmax: np.ndarray              # (p, p) float32, -1 on and below the diagonal
    stats: list


class DiscreteEngine:
    def __init__(self, p: int, n: int, max_level: int, alpha: float, max_dim: int):
        torch = nat.require_gpu()
        lib = nat.load()
        if p < 2:
            raise GpucslError("need at least two variables")
        self.torch = torch
        self.p, self.n, self.max_level, self.alpha = int(p), int(n), int(max_level), float(alpha)
        self.max_dim = int(max(max_dim, 1))
        dev = torch.device("cuda", torch.cuda.current_device())
        self.device = dev
        self.stride = (self.n + COLUMN_ALIGN - 1) // COLUMN_ALIGN * COLUMN_ALIGN
        L = max(self.max_level, 1)
        self.data = torch.zeros((p, self.stride), dtype=torch.uint8, device=dev)
        self.dims = torch.empty((p,), dtype=torch.uint8, device=dev)
        self.G = torch.empty((p, p), dtype=torch.uint8, device=dev)
        self.sep = torch.empty((p, p, L), dtype=torch.int32, device=dev)
        self.pmax = torch.empty((p, p), dtype=torch.float32, device=dev)
        self.skel16 = torch.empty((p, p), dtype=torch.int16, device=dev)
        self.stats = torch.zeros(((self.max_level + 1) * nat.STAT_SLOTS,), dtype=torch.int64, device=dev)
        self.ws_bytes = int(lib.gpucsl_disc_workspace_bytes(p, self.max_level, self.max_dim))
        self.ws = torch.empty((self.ws_bytes,), dtype=torch.uint8, device=dev)

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

    def run(self) -> None:
        nat.call("gpucsl_disc_skeleton", self.data.data_ptr(), self.n, self.stride, self.dims.data_ptr(),
                 self.p, self.max_level, self.max_dim, self.alpha, self.G.data_ptr(), self.sep.data_ptr(),
                 self.pmax.data_ptr(), self.stats.data_ptr(), self.ws.data_ptr(), self.ws_bytes,
                 nat.stream_handle())
        nat.call("gpucsl_skeleton_to_u16", self.G.data_ptr(), self.p, self.skel16.data_ptr(),
                 nat.stream_handle())

    def level_stats(self):
        return nat.decode_stats(self.stats.cpu().numpy().view(np.uint64), self.max_level)

    def results(self) -> DiscreteResult:
        skel = self.skel16.cpu().numpy().view(np.uint16).copy()
        sep = self.sep.cpu().numpy()[:, :, : self.max_level].copy()
        return DiscreteResult(skel, sep, self.pmax.cpu().numpy().copy(), self.level_stats())

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
        return table.cpu().numpy().view(np.uint32)[:cells].copy(), float(o[0]), float(o[1]), int(o[2])
