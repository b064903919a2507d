"""Device-resident Gaussian PC-stable skeleton discovery for one dataset.

GaussianEngine owns the HBM buffers of one job (correlation in FP32 and FP64,
adjacency, separation sets, zmin, pmax, statistics, workspaces) as torch
tensors and drives libgpucsl_hip.so on the current torch stream.  Nothing in
``run`` synchronises the host; ``results`` copies back.

This is the single-device part of gpucsl's GaussianDeviceManager
(gpucsl/pc/gaussian_device_manager.py:106-133, 185-193, 330-357) plus the level
loop of gaussian_ci_worker (gpucsl/pc/discover_skeleton_gaussian.py:30-113).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from gpucsl_amd import _native as nat
from gpucsl_amd.cli.cli_util import GpucslError

DEFAULT_MARGIN = 8.0
DEFAULT_GROUP_SIZE = 16


def thresholds_for_levels(thresholds, max_level: int) -> np.ndarray:
    """Per-level thresholds as float64[max_level + 1]; levels past the list get
    0.0, the reference's "never remove" value (gpucsl/pc/helpers.py:42-50)."""
    thr = np.zeros(max_level + 1, np.float64)
    k = min(len(thresholds), max_level + 1)
    thr[:k] = np.asarray(thresholds, np.float64)[:k]
    return thr


@dataclass
class GaussianResult:
    skeleton: np.ndarray       # (p, p) uint16, symmetric 0/1, zero diagonal
    separation_sets: np.ndarray  # (p, p, max_level) int32, -1 padded
    pmax: np.ndarray           # (p, p) float32, -1 on and below the diagonal
    zmin: np.ndarray           # (p, p) float32, upper triangle
    stats: list                # per-level counters


class GaussianEngine:
    def __init__(self, p: int, n: int, max_level: int, thresholds, margin: float = DEFAULT_MARGIN,
                 group_size: int = DEFAULT_GROUP_SIZE):
        torch = nat.require_gpu()
        nat.load()
        if p < 2:
            raise GpucslError("need at least two variables")
        if max_level < 0:
            raise GpucslError("max_level must be >= 0")
        self.torch = torch
        self.p, self.n, self.max_level = int(p), int(n), int(max_level)
        self.margin = float(margin)
        self.group_size = int(group_size)
        self.thr = thresholds_for_levels(thresholds, self.max_level)
        dev = torch.device("cuda", torch.cuda.current_device())
        self.device = dev
        L = max(self.max_level, 1)
        self.C32 = torch.empty((p, p), dtype=torch.float32, device=dev)
        self.C64 = torch.empty((p, p), dtype=torch.float64, device=dev)
        self.G = torch.empty((p, p), dtype=torch.uint8, device=dev)
        self.sep = torch.empty((p, p, L), dtype=torch.int32, device=dev)
        self.zmin = torch.empty((p, p), dtype=torch.float32, device=dev)
        self.pmax = torch.empty((p, p), dtype=torch.float32, device=dev)
        self.skel16 = torch.empty((p, p), dtype=torch.int16, device=dev)
        self.stats = torch.zeros(((self.max_level + 1) * nat.STAT_SLOTS,), dtype=torch.int64, device=dev)
        self.ws_bytes = int(nat.load().gpucsl_gauss_workspace_bytes(p, self.max_level))
        self.ws = torch.empty((self.ws_bytes,), dtype=torch.uint8, device=dev)
        self._corr_ws = None

    # ------------------------------------------------------------------ input
    def set_data(self, X) -> None:
        """Correlation matrix of X (n x p, float32/float64 tensor or array) on the GPU."""
        torch = self.torch
        if not isinstance(X, torch.Tensor):
            X = np.asarray(X)
            if X.dtype not in (np.float32, np.float64):
                X = X.astype(np.float64)
            X = torch.from_numpy(np.ascontiguousarray(X))
        X = X.to(self.device).contiguous()
        if X.dtype not in (torch.float32, torch.float64):
            X = X.to(torch.float64)
        n, p = X.shape
        if p != self.p:
            raise GpucslError(f"data has {p} variables, engine was built for {self.p}")
        nbytes = int(nat.load().gpucsl_corr_workspace_bytes(n, p))
        if self._corr_ws is None or self._corr_ws.numel() < nbytes:
            self._corr_ws = torch.empty((nbytes,), dtype=torch.uint8, device=self.device)
        dtype = 0 if X.dtype == torch.float32 else 1
        self._X = X  # keep alive until the stream has consumed it
        nat.call("gpucsl_corr", X.data_ptr(), dtype, n, p, self.C64.data_ptr(), self.C32.data_ptr(),
                 self._corr_ws.data_ptr(), nbytes, nat.stream_handle())

    def set_correlation(self, C) -> None:
        """Use a caller-supplied correlation matrix (host array or tensor)."""
        torch = self.torch
        if not isinstance(C, torch.Tensor):
            C = torch.from_numpy(np.ascontiguousarray(np.asarray(C, np.float64)))
        C = C.to(self.device, torch.float64).contiguous()
        if tuple(C.shape) != (self.p, self.p):
            raise GpucslError(f"correlation matrix must be {self.p}x{self.p}")
        self.C64.copy_(C)
        nat.call("gpucsl_corr_c64_to_c32", self.C64.data_ptr(), self.p, self.C32.data_ptr(),
                 nat.stream_handle())

    # -------------------------------------------------------------------- run
    def run(self) -> None:
        """All levels, back to back on the current stream, then pmax."""
        s = nat.stream_handle()
        thr = np.ascontiguousarray(self.thr)
        nat.call("gpucsl_gauss_skeleton", self.C32.data_ptr(), self.C64.data_ptr(), self.p,
                 self.max_level, thr.ctypes.data, self.margin, self.group_size, self.G.data_ptr(),
                 self.sep.data_ptr(), self.zmin.data_ptr(), self.stats.data_ptr(),
                 self.ws.data_ptr(), self.ws_bytes, s)
        self.finish()

    def run_levels(self, on_level=None) -> None:
        """Level by level (for per-level logging); on_level(level) after each."""
        s = nat.stream_handle()
        nat.call("gpucsl_gauss_init", self.p, self.max_level, self.G.data_ptr(), self.sep.data_ptr(),
                 self.zmin.data_ptr(), self.stats.data_ptr(), s)
        nat.call("gpucsl_gauss_level0", self.C32.data_ptr(), self.C64.data_ptr(), self.p,
                 float(self.thr[0]), self.margin, self.G.data_ptr(), self.zmin.data_ptr(),
                 self.stats.data_ptr(), self.ws.data_ptr(), self.ws_bytes, s)
        if on_level:
            on_level(0)
        for l in range(1, self.max_level + 1):
            st = self.stats.data_ptr() + 8 * nat.STAT_SLOTS * l
            nat.call("gpucsl_gauss_level", l, self.C32.data_ptr(), self.C64.data_ptr(), self.p,
                     self.max_level, float(self.thr[l]), self.margin, self.group_size,
                     self.G.data_ptr(), self.sep.data_ptr(), self.zmin.data_ptr(), st,
                     self.ws.data_ptr(), self.ws_bytes, s)
            if on_level:
                on_level(l)
        self.finish()

    def capture(self, X=None) -> None:
        """Capture set_data(X) (when X is given) + run() into one HIP graph;
        replay() then repeats the whole skeleton with one launch.  Buffers
        and X stay bound to the graph."""
        torch = self.torch
        if X is not None:
            self.set_data(X)
        self.run()
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            g.capture_begin()
            if X is not None:
                self.set_data(X)
            self.run()
            g.capture_end()
        torch.cuda.current_stream().wait_stream(s)
        self._graph = g

    def replay(self) -> None:
        self._graph.replay()

    def finish(self) -> None:
        s = nat.stream_handle()
        nat.call("gpucsl_gauss_outputs", self.zmin.data_ptr(), self.G.data_ptr(), self.p, self.n,
                 self.pmax.data_ptr(), self.skel16.data_ptr(), s)

    # ----------------------------------------------------------------- output
    def level_stats(self):
        return nat.decode_stats(self.stats.cpu().numpy().view(np.uint64), self.max_level)

    def results(self) -> GaussianResult:
        skel, sep, pmax, zmin, st = nat.to_host(self.skel16, self.sep, self.pmax, self.zmin, self.stats)
        return GaussianResult(skel.view(np.uint16), sep[:, :, : self.max_level], pmax, zmin,
                              nat.decode_stats(st.view(np.uint64), self.max_level))
