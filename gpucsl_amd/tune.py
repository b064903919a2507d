"""Autotuner for the level-l CI-test kernel's launch shapes.

The library keeps one shape per (launch bucket, level): one dataset with p <=
256, one dataset with p > 256, and a batch of several datasets
(GPUCSL_TUNING_SMALL / _LARGE / _BATCH in include/gpucsl_hip.h); the degree of
a row is handled on the device, where each row's subset count C(deg - 1, l)
picks its team size against the shape's thresholds.

For a bucket's workload (a single synthetic dataset, or the batched workload's
bootstrap resamples and uneven datasets) the tuner runs a staged search:
starting from the current shapes it varies one field at a time over its
values with the other fields fixed, keeps per level the value with the least
time of that level's kernels (library launch events), and repeats the pass
over the fields.  One run times every level, so each candidate value costs
one run for all levels.  Shapes never change results (the tests run in
subset-rank order and the first passing one wins whatever the team size); the
sweep checks that and only measures time.  The table is written as JSON;
``apply_table`` loads it into the library, and ``_native.load`` applies
``gpucsl_amd/tuning.json`` when that file exists.

    python -m gpucsl_amd.tune --bucket large --p 1000 --n 10000 --degree 4 --max-level 4 --out t_large.json
    python -m gpucsl_amd.tune --bucket batch --out t_batch.json
    python -m gpucsl_amd.tune --merge t_large.json t_batch.json t_small.json --out gpucsl_amd/tuning.json
"""
from __future__ import annotations

import argparse
import ctypes as ct
import json
import sys
import time
from pathlib import Path

import numpy as np

from gpucsl_amd import _native as nat

SPACE = {
    "small_team_iters": (1, 2, 4, 8, 16, 32),
    "wave_team_iters": (2, 4, 8, 16),
    "fast_blocks_per_cu": (2, 4, 8, 16, 24, 32),
    "fast_lds_bytes": (4096, 6144, 8192, 16384, 32768, 65536),
    "item_tests": (256, 1024, 4096, 16384),
    "wave_items": (0, 1),
}
FIELDS = ("small_team_iters", "wave_team_iters", "fast_blocks_per_cu", "fast_lds_bytes", "item_tests",
          "wave_items")
BUCKETS = {"small": 0, "large": 1, "batch": 2}   # GPUCSL_TUNING_SMALL / _LARGE / _BATCH
MAX_TUNED_LEVEL = 8                              # GPUCSL_MAX_LEVEL_TEMPLATE


class LevelTuning(ct.Structure):
    _fields_ = [(f, ct.c_int32) for f in FIELDS]


def set_shape(bucket: int, level: int, shape: dict) -> None:
    """bucket -1: every bucket; level 0: every level."""
    t = LevelTuning(*[int(shape[f]) for f in FIELDS])
    nat.call("gpucsl_set_bucket_tuning", bucket, level, ct.byref(t))


def get_shape(bucket: int, level: int) -> dict:
    t = LevelTuning()
    nat.call("gpucsl_get_bucket_tuning", bucket, level, ct.byref(t))
    return {f: int(getattr(t, f)) for f in FIELDS}


def _apply_levels(bucket: int, levels: dict) -> None:
    buckets = range(len(BUCKETS)) if bucket < 0 else (bucket,)
    for lv, shape in levels.items():
        for b in buckets:
            cur = get_shape(b, int(lv))
            cur.update({f: shape[f] for f in FIELDS if f in shape})
            set_shape(b, int(lv), cur)


def apply_table(table: dict) -> None:
    """Apply a tuning table: "buckets" maps a bucket name to its "levels";
    a top-level "levels" applies to every bucket.  Fields a table lacks keep
    their current values."""
    if "levels" in table:
        _apply_levels(-1, table["levels"])
    for name, sub in table.get("buckets", {}).items():
        _apply_levels(BUCKETS[name], sub.get("levels", {}))


def _level_ms(timings: dict, level: int, reps: int) -> float:
    tot = 0.0
    for name in (f"gauss_level{level}_fast", f"gauss_level{level}_slow"):
        if name in timings:
            tot += timings[name][1]
    return tot / reps


class _Workload:
    """A bucket's workload: run() launches the skeleton(s) eagerly, result()
    gives what must not change."""

    def __init__(self, bucket: str, p: int, n: int, degree: float, max_level: int, alpha: float, seed: int):
        import torch

        from gpucsl_amd import synthetic as syn

        self.torch = torch
        self.max_level = max_level
        if bucket == "batch":
            from gpucsl_amd.batch import GaussJobState, GaussianBatch

            base, rows, sets = syn.batched_workload(256, 64)
            base_dev = torch.from_numpy(base).cuda()
            jobs = []
            for r in rows:
                j = GaussJobState(5000, 200, max_level, alpha=alpha)
                j.set_data(base_dev[torch.from_numpy(r).cuda()])
                jobs.append(j)
            for X in sets:
                j = GaussJobState(X.shape[0], X.shape[1], max_level, alpha=alpha)
                j.set_data(X)
                jobs.append(j)
            self.batch = GaussianBatch(jobs)
            self.desc = {"jobs": len(jobs), "workload": "256 bootstrap resamples of p=200 n=5000 + 64 datasets "
                                                       "p~U[50,500] n~U[1000,20000]", "max_level": max_level}
        else:
            from gpucsl_amd.gaussian import GaussianEngine
            from gpucsl_amd.pc.helpers import get_gaussian_thresholds

            data = syn.gaussian_sem(p, n, degree, seed=seed).data.astype(np.float32)
            self.eng = GaussianEngine(p, n, max_level, get_gaussian_thresholds(data, alpha))
            self.eng.set_data(torch.from_numpy(data).cuda())
            self.batch = None
            self.desc = {"p": p, "n": n, "degree": degree, "max_level": max_level, "alpha": alpha, "seed": seed}

    def run(self) -> None:
        if self.batch is not None:
            self.batch.launch()
        else:
            self.eng.run()

    def result(self):
        self.torch.cuda.synchronize()
        if self.batch is not None:
            return [(r.skeleton.copy(), r.separation_sets.copy()) for r in self.batch.results()]
        r = self.eng.results()
        return [(r.skeleton, r.separation_sets)]


def staged_search(bucket: str, work: _Workload, reps: int = 3, rounds: int = 2, log=print) -> dict:
    """Per level, the shape found by varying one field at a time (see the
    module docstring); returns {level: (ms, shape)}."""
    torch = work.torch
    b = BUCKETS[bucket]
    levels = range(1, min(work.max_level, MAX_TUNED_LEVEL) + 1)
    work.run()
    want = work.result()
    best = {lv: [get_shape(b, lv), float("inf")] for lv in levels}

    def measure(candidate: dict) -> dict:
        for lv in levels:
            set_shape(b, lv, candidate[lv])
        work.run()
        torch.cuda.synchronize()
        with nat.KernelTimer() as kt:
            for _ in range(reps):
                work.run()
            torch.cuda.synchronize()
        got = work.result()
        if any(not ((g[0] == w[0]).all() and (g[1] == w[1]).all()) for g, w in zip(got, want)):
            raise RuntimeError(f"shape {candidate} changed the result")
        return {lv: _level_ms(kt.timings, lv, reps) for lv in levels}

    base = measure({lv: best[lv][0] for lv in levels})
    for lv in levels:
        best[lv][1] = base[lv]
    evals, t0 = 1, time.perf_counter()
    for rnd in range(rounds):
        for field in FIELDS:
            for value in SPACE[field]:
                cand = {lv: dict(best[lv][0], **{field: value}) for lv in levels}
                ms = measure(cand)
                evals += 1
                for lv in levels:
                    if ms[lv] < best[lv][1]:
                        best[lv] = [cand[lv], ms[lv]]
            log(f"[tune] {bucket} round {rnd + 1} {field}: "
                + ", ".join(f"L{lv} {best[lv][1]:.4f} ms" for lv in levels)
                + f" ({evals} runs, {time.perf_counter() - t0:.1f}s)")
    for lv in levels:
        set_shape(b, lv, best[lv][0])
    return {lv: (best[lv][1], best[lv][0]) for lv in levels}


def sweep(bucket: str, p: int, n: int, degree: float, max_level: int, alpha: float = 0.01, seed: int = 100,
          reps: int = 3, rounds: int = 2, log=print) -> dict:
    import torch

    work = _Workload(bucket, p, n, degree, max_level, alpha, seed)
    best = staged_search(bucket, work, reps=reps, rounds=rounds, log=log)
    return {"device": torch.cuda.get_device_name(0),
            "buckets": {bucket: {"workload": work.desc,
                                 "levels": {str(lv): dict(shape, ms=round(ms, 5)) for lv, (ms, shape) in best.items()}}}}


def merge(tables: list) -> dict:
    out = {"buckets": {}}
    for t in tables:
        out["device"] = t.get("device", out.get("device"))
        out["buckets"].update(t.get("buckets", {}))
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--bucket", choices=sorted(BUCKETS), default="large")
    ap.add_argument("--p", type=int, default=1000)
    ap.add_argument("--n", type=int, default=10000)
    ap.add_argument("--degree", type=float, default=4.0)
    ap.add_argument("--max-level", type=int, default=None, help="default 4 (3 for the batch bucket)")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--rounds", type=int, default=2)
    ap.add_argument("--merge", nargs="+", metavar="TABLE", help="merge bucket tables into one")
    ap.add_argument("--out", default="tuning.json")
    a = ap.parse_args(argv)
    if a.merge:
        table = merge([json.loads(Path(f).read_text()) for f in a.merge])
    else:
        nat.require_gpu()
        L = a.max_level if a.max_level is not None else (3 if a.bucket == "batch" else 4)
        table = sweep(a.bucket, a.p, a.n, a.degree, L, reps=a.reps, rounds=a.rounds,
                      log=lambda *x: print(*x, file=sys.stderr, flush=True))
    Path(a.out).write_text(json.dumps(table, indent=1))
    print(json.dumps(table))
    return 0


if __name__ == "__main__":
    sys.exit(main())
