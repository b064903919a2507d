"""Continuous-batching server for skeleton-discovery requests on one GPU.

Requests (Gaussian or discrete datasets) arrive at any time.  The server
admits them in arrival order while their HBM footprint fits its budget, keeps
every admitted job's state on the device (data, correlation, adjacency,
separation sets, pmax, level workspace) and advances the jobs one level per
scheduling round ("tick"):

- Gaussian jobs that sit at the same stage share one batched launch per
  kernel (stage 0 = correlation + level 0, stage l = level l, last stage =
  pmax/skeleton), so a job that arrives while others are at level 2 starts its
  level 0 in the same round as their level 3;
- discrete jobs are spread over their own streams (round robin at
  admission) and advance the same way: the jobs of one stream at the same
  stage share one batched launch per kernel (gpucsl_disc_batch_*), and the
  streams overlap with the Gaussian batches and with each other;
- at most ``admit_per_tick`` requests are admitted per round, so a burst of
  arrivals runs as groups at different stages.

Results are copied to pinned host memory and handed back when their copy has
completed; latency is measured from ``submit`` to that point.  A job's result
does not depend on the other jobs in flight: the Gaussian kernels give the
same bits in any batch (the correlation split-K is a function of the job's own
n and p), and so do the discrete ones (exact integer tables, the same per-test
arithmetic, removal keys and pmax by atomic min and max).

This replaces calling gpucsl's discover_skeleton_gpu_gaussian / _discrete
(gpucsl/pc/discover_skeleton_gaussian.py:116-172,
gpucsl/pc/discover_skeleton_discrete.py:24-128) once per request.
"""
from __future__ import annotations

import ctypes as ct
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from gpucsl_amd import _native as nat
from gpucsl_amd.batch import GaussJob, GaussJobState, job_footprint
from gpucsl_amd.cli.cli_util import GpucslError
from gpucsl_amd.discrete import DiscJob, DiscreteEngine, DiscreteResult, disc_layout
from gpucsl_amd.gaussian import DEFAULT_GROUP_SIZE, DEFAULT_MARGIN, GaussianResult

MAX_SERVED_LEVEL = 32    # GPUCSL_MAX_LEVEL: the deepest level any kernel runs

# The server's streams, created once per process and device and shared by
# every server: HIP gives each new stream the next of its hardware queues in
# turn (four per process by default), and streams that share a queue run in
# submission order, so fresh streams for every server would land on queues
# that other streams already use.  Roles: "g" Gaussian batches, ("d", k)
# discrete batches.
_STREAMS: dict = {}


def _stream(torch, role):
    key = (torch.cuda.current_device(), role)
    if key not in _STREAMS:
        _STREAMS[key] = torch.cuda.Stream()
    return _STREAMS[key]


@dataclass
class Request:
    kind: str                  # "gaussian" or "discrete"
    n: int
    p: int
    max_level: int = 3
    alpha: float = 0.01
    X: object = None           # gaussian: (n, p) float32 or float64 tensor or array (device or host)
    codes: object = None       # discrete: (p, >= n) uint8 device tensor, one row per variable
    dims: object = None        # discrete: (p,) category counts
    max_dim: int = 0           # discrete: largest category count (0 = from dims)
    rid: int = -1
    arrival: float = 0.0


@dataclass
class Completed:
    rid: int
    kind: str
    result: object             # GaussianResult or DiscreteResult
    arrival: float
    done: float
    stats: list = field(default_factory=list)
    _release: object = None    # zero-copy results: gives the pinned buffers back

    @property
    def latency(self) -> float:
        return self.done - self.arrival

    def release(self) -> None:
        """Zero-copy and device results only: the arrays are views of the
        server's pinned pool (or, with results on the device, tensors in its
        device pool) and stay valid until this call."""
        if self._release is not None:
            self._release()
            self._release = None
            self.result = None


def x_dtype_of(req: Request) -> str:
    """The dtype a Gaussian request's correlation is computed from: float64 X
    stays float64 (as in the single-dataset path); anything else is float32."""
    dt = str(getattr(req.X, "dtype", "float32"))
    return "float64" if dt in ("float64", "torch.float64") else "float32"


def request_footprint(req: Request) -> int:
    """HBM bytes of a request's device state (data, matrices, separation sets,
    workspaces), the unit of admission."""
    if req.kind == "gaussian":
        return job_footprint(req.n, req.p, req.max_level, 8 if x_dtype_of(req) == "float64" else 4)
    if req.kind == "discrete":
        return disc_layout(req.n, req.p, req.max_level, max(req.max_dim, 2))[1]
    raise GpucslError(f"unknown request kind {req.kind!r}")


def result_bytes(req: Request) -> int:
    """Pinned host bytes of a request's results (skeleton, sepsets, pmax, and
    zmin or the statistics)."""
    pp, L = req.p * req.p, max(req.max_level, 1)
    return _al(2 * pp) + _al(4 * pp * L) + _al(4 * pp) + _al(max(4 * pp, 8 * (MAX_SERVED_LEVEL + 1) * 8))


def _al(x: int) -> int:
    return (x + 255) // 256 * 256


class Pool:
    """First-fit sub-allocator over one preallocated buffer: offsets are
    256-byte aligned, freed ranges coalesce with their neighbours."""

    def __init__(self, size: int):
        self.size = int(size)
        self.free = [(0, self.size)]      # sorted (offset, length)
        self.in_use = 0

    def alloc(self, nbytes: int):
        need = _al(max(int(nbytes), 1))
        for k, (off, ln) in enumerate(self.free):
            if ln >= need:
                if ln == need:
                    self.free.pop(k)
                else:
                    self.free[k] = (off + need, ln - need)
                self.in_use += need
                return off, need
        return None

    def release(self, off: int, ln: int) -> None:
        self.in_use -= ln
        fr = self.free
        k = 0
        while k < len(fr) and fr[k][0] < off:
            k += 1
        fr.insert(k, (off, ln))
        if k + 1 < len(fr) and fr[k][0] + fr[k][1] == fr[k + 1][0]:
            fr[k] = (fr[k][0], fr[k][1] + fr[k + 1][1])
            fr.pop(k + 1)
        if k > 0 and fr[k - 1][0] + fr[k - 1][1] == fr[k][0]:
            fr[k - 1] = (fr[k - 1][0], fr[k - 1][1] + fr[k][1])
            fr.pop(k)


class _Job:
    __slots__ = ("req", "stage", "state", "stream", "dev_range", "host_range", "desc", "host", "done_event", "last")

    def __init__(self, req, state, stream, dev_range, host_range):
        self.req, self.state, self.stream = req, state, stream
        self.dev_range, self.host_range = dev_range, host_range
        self.stage = 0
        self.desc = None
        self.host = None
        self.done_event = None
        self.last = req.max_level + 1      # the finish stage


class SkeletonServer:
    """hbm_budget: bytes of the device pool that holds every admitted job's
    state (allocated once, up front); host_budget: bytes of pinned host memory
    for results in transit."""

    def __init__(self, hbm_budget: int | None = None, host_budget: int | None = None,
                 margin: float = DEFAULT_MARGIN, group_size: int = DEFAULT_GROUP_SIZE, discrete_streams: int = 2,
                 ticks_in_flight: int = 2, copy_results: bool = True, results_on_device: bool = False,
                 admit_per_tick: int | None = 8):
        torch = nat.require_gpu()
        self.torch = torch
        self.lib = nat.load()
        self.margin, self.group_size = float(margin), int(group_size)
        if hbm_budget is None:
            free, _ = torch.cuda.mem_get_info()
            hbm_budget = min(int(0.5 * free), 32 << 30)
        self.budget = _al(int(hbm_budget))
        self.host_budget = _al(int(host_budget if host_budget is not None else min(self.budget, 2 << 30)))
        self.dev_pool = Pool(self.budget)
        self.host_pool = Pool(self.host_budget)
        self.dev_mem = torch.empty((self.budget,), dtype=torch.uint8, device="cuda")
        self.host_mem = torch.empty((self.host_budget,), dtype=torch.uint8, pin_memory=True)
        self._ws = None
        # copy_results=False hands out views of the pinned pool (no host copy);
        # the caller gives them back with Completed.release().  With
        # results_on_device the skeleton, sepsets, pmax and zmin stay in the
        # job's device range (torch tensors, nothing crosses to the host but
        # the discrete statistics) until Completed.release().
        self.copy_results = bool(copy_results)
        self.results_on_device = bool(results_on_device)
        self.gstream = _stream(torch, "g")
        self.dstreams = [_stream(torch, ("d", k)) for k in range(max(1, discrete_streams))]
        self._rr = 0
        self.pending: deque = deque()
        self.active: list = []
        # per stream: launched rounds not yet complete, as (event, keep-alive)
        self.outstanding = {st: deque() for st in [self.gstream] + self.dstreams}
        self.ticks_in_flight = max(1, int(ticks_in_flight))
        # at most this many admissions per round (None: as many as fit): a
        # burst of arrivals then runs as overlapping groups at different
        # stages, whose batched launches keep the device busier than one
        # group moving stage by stage (the 96-request mix, all at t = 0:
        # 64 -> 59 ms with results left on the device)
        self.admit_per_tick = None if admit_per_tick is None else max(1, int(admit_per_tick))
        self.completed: list = []
        self._next_id = 0
        with torch.cuda.stream(self.gstream):
            # Gaussian statistics, accumulated over every batched launch
            self.gstats = torch.zeros(((MAX_SERVED_LEVEL + 1) * nat.STAT_SLOTS,), dtype=torch.int64,
                                      device="cuda")
        self.ticks = 0
        self.host_time = {"reap": 0.0, "admit": 0.0, "launch": 0.0}   # host seconds per tick phase
        self.launched = False
        self.max_batch = 0

    # ------------------------------------------------------------- requests
    def submit(self, req: Request, arrival: float | None = None) -> int:
        if req.kind not in ("gaussian", "discrete"):
            raise GpucslError(f"unknown request kind {req.kind!r}")
        if not 0 <= req.max_level <= MAX_SERVED_LEVEL:
            raise GpucslError(f"max_level must be in 0..{MAX_SERVED_LEVEL}")
        if req.kind == "gaussian" and (req.X is None or tuple(req.X.shape) != (req.n, req.p)):
            raise GpucslError("a Gaussian request needs X of shape (n, p)")
        if req.kind == "discrete":
            if req.codes is None or req.dims is None:
                raise GpucslError("a discrete request needs codes and dims")
            if req.max_dim <= 0:
                req.max_dim = int(req.dims.max().item()) if hasattr(req.dims, "max") else int(np.max(req.dims))
        fp = request_footprint(req)
        if fp > self.budget:
            raise GpucslError(f"request needs {fp} bytes of HBM, more than the budget {self.budget}")
        if not (self.copy_results and not self.results_on_device) and result_bytes(req) > self.host_budget:
            raise GpucslError(f"request results need {result_bytes(req)} bytes, more than the host budget")
        req.rid = self._next_id
        self._next_id += 1
        req.arrival = time.perf_counter() if arrival is None else arrival
        self.pending.append(req)
        return req.rid

    @property
    def idle(self) -> bool:
        return not self.pending and not self.active

    # ------------------------------------------------------------ scheduling
    def tick(self) -> list:
        """One scheduling round: collect finished requests, admit, and launch
        the next stage of every job whose stream has fewer than
        ``ticks_in_flight`` rounds outstanding.  Never blocks; returns the
        requests completed since the last round."""
        t0 = time.perf_counter()
        done = self._reap()
        t1 = time.perf_counter()
        self._admit()
        t2 = time.perf_counter()
        self.launched = self._launch()
        t3 = time.perf_counter()
        ht = self.host_time
        ht["reap"] += t1 - t0
        ht["admit"] += t2 - t1
        ht["launch"] += t3 - t2
        self.ticks += 1
        return done

    def wait(self, timeout: float = 0.05) -> None:
        """Sleep until some outstanding round or result copy completes (or the
        timeout passes)."""
        events = [q[0][0] for q in self.outstanding.values() if q]
        events += [j.done_event for j in self.active if j.done_event is not None]
        if not events:
            return
        end = time.perf_counter() + timeout
        while not any(e.query() for e in events) and time.perf_counter() < end:
            time.sleep(5e-5)

    def drain(self) -> list:
        out = []
        while not self.idle:
            got = self.tick()
            out.extend(got)
            if not got and not self.launched:
                self.wait()
        return out

    @property
    def used(self) -> int:
        return self.dev_pool.in_use

    def _admit(self) -> None:
        """Admit in arrival order while the head request's state fits the
        device pool and its results fit the host pool."""
        torch = self.torch
        admitted = 0
        while self.pending and (self.admit_per_tick is None or admitted < self.admit_per_tick):
            admitted += 1
            req = self.pending[0]
            dr = self.dev_pool.alloc(request_footprint(req))
            if dr is None:
                break
            # owned results get pinned tensors of their own at the finish
            # stage; zero-copy and device results use the pinned pool
            owned = self.copy_results and not self.results_on_device
            hr = None if owned else self.host_pool.alloc(result_bytes(req))
            if hr is None and not owned:
                self.dev_pool.release(*dr)
                break
            self.pending.popleft()
            arena = self.dev_mem[dr[0]:dr[0] + dr[1]]
            if req.kind == "gaussian":
                with torch.cuda.stream(self.gstream):
                    st = GaussJobState(req.n, req.p, req.max_level, alpha=req.alpha, arena=arena,
                                       x_dtype=x_dtype_of(req))
                    X = req.X
                    if getattr(X, "is_cuda", False):
                        X.record_stream(self.gstream)
                    st.set_data(X)
                job = _Job(req, st, self.gstream, dr, hr)
                job.desc = bytes(st.descriptor())
            else:
                s = self.dstreams[self._rr % len(self.dstreams)]
                self._rr += 1
                with torch.cuda.stream(s):
                    eng = DiscreteEngine(req.p, req.n, req.max_level, req.alpha, req.max_dim, arena=arena)
                    codes, dims = req.codes, req.dims
                    if not isinstance(codes, torch.Tensor):
                        codes = torch.from_numpy(np.ascontiguousarray(codes)).cuda()
                    if not isinstance(dims, torch.Tensor):
                        dims = torch.from_numpy(np.ascontiguousarray(dims, np.uint8)).cuda()
                    codes.record_stream(s)
                    dims.record_stream(s)
                    eng.set_codes_device(codes, dims)
                job = _Job(req, eng, s, dr, hr)
            self.active.append(job)

    def _host_views(self, j, torch):
        """Pinned result buffers of a job: inside the host pool, or (owned
        results) pinned tensors of its own that the caller keeps."""
        p, L = j.req.p, max(j.req.max_level, 1)
        if j.host_range is None:
            out = {"skeleton": torch.empty((p, p), dtype=torch.int16, pin_memory=True),
                   "sepsets": torch.empty((p, p, L), dtype=torch.int32, pin_memory=True),
                   "pmax": torch.empty((p, p), dtype=torch.float32, pin_memory=True)}
            if j.req.kind == "gaussian":
                out["zmin"] = torch.empty((p, p), dtype=torch.float32, pin_memory=True)
            else:
                out["stats"] = torch.empty(((j.req.max_level + 1) * nat.STAT_SLOTS,), dtype=torch.int64,
                                           pin_memory=True)
            return out
        base, pp = j.host_range[0], p * p
        sizes = [("skeleton", torch.int16, (p, p), 2 * pp), ("sepsets", torch.int32, (p, p, L), 4 * pp * L),
                 ("pmax", torch.float32, (p, p), 4 * pp)]
        if j.req.kind == "gaussian":
            sizes.append(("zmin", torch.float32, (p, p), 4 * pp))
        else:
            n_st = (j.req.max_level + 1) * nat.STAT_SLOTS
            sizes.append(("stats", torch.int64, (n_st,), 8 * n_st))
        out, off = {}, base
        for name, dt, shape, nb in sizes:
            out[name] = self.host_mem[off:off + nb].view(dt).view(*shape)
            off += _al(nb)
        return out

    def _batch_ws(self, nbytes: int):
        """One launch workspace for every batched launch: they all run on the
        Gaussian stream, in order, so they can share it; it only grows."""
        if self._ws is None or self._ws.numel() < nbytes:
            with self.torch.cuda.stream(self.gstream):
                self._ws = self.torch.empty((max(nbytes, 1 << 20),), dtype=self.torch.uint8, device="cuda")
        return self._ws

    def _launch(self) -> bool:
        torch = self.torch
        ready = {st for st, q in self.outstanding.items() if len(q) < self.ticks_in_flight}
        used, keep = set(), []
        groups: dict = {}
        dgroups: dict = {}
        for j in self.active:
            if j.done_event is not None or j.stream not in ready:
                continue
            if j.req.kind == "gaussian":
                groups.setdefault(j.stage, []).append(j)
            else:
                key = (self.dstreams.index(j.stream), j.stage, j.stage == j.last)
                dgroups.setdefault(key, []).append(j)
        for (si, stage, finish), jobs in sorted(dgroups.items()):
            self._launch_discrete_group(jobs, stage, finish, self.dstreams[si])
            used.add(self.dstreams[si])
        if groups:
            used.add(self.gstream)
            with torch.cuda.stream(self.gstream):
                s = nat.stream_handle()
                for stage in sorted(groups):
                    self._launch_gaussian_group(groups[stage], stage, s, keep)
        for st in used:
            e = torch.cuda.Event()
            e.record(st)
            self.outstanding[st].append((e, keep if st is self.gstream else None))
        return bool(used)

    def _table(self, jobs, keep):
        torch = self.torch
        raw = b"".join(j.desc for j in jobs)
        host = (GaussJob * len(jobs)).from_buffer_copy(raw)
        pinned = torch.empty((len(raw),), dtype=torch.uint8, pin_memory=True)
        pinned.numpy()[:] = np.frombuffer(raw, np.uint8)
        dev = torch.empty((len(raw),), dtype=torch.uint8, device="cuda")
        dev.copy_(pinned, non_blocking=True)
        keep += [host, pinned, dev]
        return host, dev

    def _launch_gaussian_group(self, jobs, stage, s, keep) -> None:
        torch = self.torch
        # jobs at the finish stage and jobs at a level share nothing: split them
        finishing = [j for j in jobs if j.stage == j.last]
        running = [j for j in jobs if j.stage < j.last]
        if running:
            host, dev = self._table(running, keep)
            nj = len(running)
            self.max_batch = max(self.max_batch, nj)
            st = self.gstats.data_ptr()
            if stage == 0:
                nat.call("gpucsl_gauss_batch_corr", ct.byref(host), dev.data_ptr(), nj, s)
                nat.call("gpucsl_gauss_batch_init", ct.byref(host), dev.data_ptr(), nj, s)
                nat.call("gpucsl_gauss_batch_level0", ct.byref(host), dev.data_ptr(), nj, self.margin, st, s)
            else:
                ws_bytes = int(self.lib.gpucsl_gauss_batch_workspace_bytes(ct.byref(host), nj))
                ws = self._batch_ws(ws_bytes)
                nat.call("gpucsl_gauss_batch_level", stage, ct.byref(host), dev.data_ptr(), nj, self.margin,
                         self.group_size, st + 8 * nat.STAT_SLOTS * stage, ws.data_ptr(), ws_bytes, s)
            for j in running:
                j.stage += 1
        if finishing:
            host, dev = self._table(finishing, keep)
            nat.call("gpucsl_gauss_batch_finish", ct.byref(host), dev.data_ptr(), len(finishing), s)
            for j in finishing:
                self._copy_out_gaussian(j)

    def _copy_out_gaussian(self, j) -> None:
        torch = self.torch
        st: GaussJobState = j.state
        p, L = st.p, max(st.max_level, 1)
        if self.results_on_device:
            j.host = {"skeleton": st.view("skeleton", torch.int16, (p, p)),
                      "sepsets": st.view("sepsets", torch.int32, (p, p, L)),
                      "pmax": st.view("pmax", torch.float32, (p, p)),
                      "zmin": st.view("zmin", torch.float32, (p, p))}
            j.done_event = torch.cuda.Event()
            j.done_event.record(self.gstream)
            return
        # the copies stay on the Gaussian stream: on a stream of their own
        # (waiting on an event of this one) the closed loop measured 105 ->
        # 140 ms for the served mix
        h = self._host_views(j, torch)
        h["skeleton"].copy_(st.view("skeleton", torch.int16, (p, p)), non_blocking=True)
        h["sepsets"].copy_(st.view("sepsets", torch.int32, (p, p, L)), non_blocking=True)
        h["pmax"].copy_(st.view("pmax", torch.float32, (p, p)), non_blocking=True)
        h["zmin"].copy_(st.view("zmin", torch.float32, (p, p)), non_blocking=True)
        j.host = h
        j.done_event = torch.cuda.Event()
        j.done_event.record(self.gstream)

    def _launch_discrete_group(self, jobs, stage, finish, stream) -> None:
        """One stage of discrete jobs that share a stream: set-up and level 0,
        a level, or (finish) their finish and the result copies, batched."""
        torch = self.torch
        table = (DiscJob * len(jobs))(*[j.state.descriptor() for j in jobs])
        self.max_batch = max(self.max_batch, len(jobs))
        with torch.cuda.stream(stream):
            s = nat.stream_handle()
            if finish:
                nat.call("gpucsl_disc_batch_finish", ct.byref(table), len(jobs), s)
            elif stage == 0:
                nat.call("gpucsl_disc_batch_init", ct.byref(table), len(jobs), s)
                nat.call("gpucsl_disc_batch_level0", ct.byref(table), len(jobs), s)
            else:
                nat.call("gpucsl_disc_batch_level", stage, ct.byref(table), len(jobs), s)
            for j in jobs:
                if not finish:
                    j.stage += 1
                    continue
                eng: DiscreteEngine = j.state
                h = self._host_views(j, torch)
                h["stats"].copy_(eng.stats, non_blocking=True)
                if self.results_on_device:
                    h["skeleton"], h["sepsets"], h["pmax"] = eng.skel16, eng.sep, eng.pmax
                else:
                    h["skeleton"].copy_(eng.skel16, non_blocking=True)
                    h["sepsets"].copy_(eng.sep, non_blocking=True)
                    h["pmax"].copy_(eng.pmax, non_blocking=True)
                j.host = h
                j.done_event = torch.cuda.Event()
                j.done_event.record(stream)

    def _reap(self) -> list:
        """Jobs whose device work and result copy are complete give back their
        device range at once and complete: owned results are the pinned
        tensors their copy went into, zero-copy results are views of the
        pinned pool until released.  With results on the device a finished job
        completes at once and keeps its device range until the caller
        releases the result."""
        still, zero_copy = [], []
        for j in self.active:
            if j.done_event is not None and j.done_event.query():
                if self.results_on_device:
                    # the results are the job's device state: both ranges go
                    # back when the caller releases them
                    c = self._result(j, copy=False)
                    c._release = (lambda dr=j.dev_range, hr=j.host_range:
                                  (self.dev_pool.release(*dr), self.host_pool.release(*hr)))
                    zero_copy.append(c)
                    continue
                self.dev_pool.release(*j.dev_range)
                if j.host_range is None:
                    # owned results: the pinned tensors are the caller's
                    zero_copy.append(self._result(j, copy=False))
                else:
                    c = self._result(j, copy=False)
                    c._release = (lambda hr=j.host_range: self.host_pool.release(*hr))
                    zero_copy.append(c)
            else:
                still.append(j)
        self.active = still
        out = zero_copy
        # release launch tables and workspaces of completed rounds
        for q in self.outstanding.values():
            while q and q[0][0].query():
                q.popleft()
        self.completed.extend(out)
        return out

    def _result(self, j, copy: bool = True) -> Completed:
        req, h = j.req, j.host
        L = req.max_level
        if self.results_on_device:
            # torch tensors on the device; the skeleton as int16 (uint16 bits)
            skel, sep, pmax = h["skeleton"], h["sepsets"][:, :, :L], h["pmax"]
            zmin = h.get("zmin")
        else:
            own = (lambda a: a.copy()) if copy else (lambda a: a)
            skel = own(h["skeleton"].numpy().view(np.uint16))
            sep = own(h["sepsets"].numpy()[:, :, :L])
            pmax = own(h["pmax"].numpy())
            zmin = own(h["zmin"].numpy()) if req.kind == "gaussian" else None
        if req.kind == "gaussian":
            res = GaussianResult(skel, sep, pmax, zmin, [])
            stats = []
        else:
            stats = nat.decode_stats(h["stats"].numpy().view(np.uint64), L)
            res = DiscreteResult(skel, sep, pmax, stats)
        return Completed(req.rid, req.kind, res, req.arrival, time.perf_counter(), stats)

    # ------------------------------------------------------------ statistics
    def gaussian_stats(self):
        self.torch.cuda.synchronize()
        return nat.decode_stats(self.gstats.cpu().numpy().view(np.uint64), MAX_SERVED_LEVEL)

    def useful_tests(self) -> int:
        g = sum(d["useful"] for d in self.gaussian_stats())
        d = sum(sum(s["useful"] for s in c.stats) for c in self.completed if c.kind == "discrete")
        return g + d


def run_alone(req: Request, margin: float = DEFAULT_MARGIN, group_size: int = DEFAULT_GROUP_SIZE):
    """One request through the single-dataset path on the current stream (the
    comparison for the served results and throughput)."""
    from gpucsl_amd.batch import run_plain

    torch = nat.require_gpu()
    if req.kind == "gaussian":
        st = GaussJobState(req.n, req.p, req.max_level, alpha=req.alpha, x_dtype=x_dtype_of(req))
        st.set_data(req.X)
        run_plain([st], margin, group_size)
        return st
    max_dim = req.max_dim or int(req.dims.max().item())
    eng = DiscreteEngine(req.p, req.n, req.max_level, req.alpha, max_dim)
    eng.set_codes_device(req.codes, req.dims if isinstance(req.dims, torch.Tensor)
                         else torch.from_numpy(np.asarray(req.dims, np.uint8)).cuda())
    eng.run()
    return eng
