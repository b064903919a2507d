"""Autotuner for the level-l CI-test kernel's launch shapes.

For a workload (synthetic data of a given p, n and degree) it runs the
skeleton once per candidate shape, times every level's kernels with the
library's launch events, and keeps, per level, the shape with the least time.
Shapes never change results (the tests run in subset-rank order and the first
passing one wins whatever the team size), so the sweep only measures time.
The table is written as JSON; ``apply_table`` loads it into the library, and
``_native.load`` applies ``gpucsl_amd/tuning.json`` when that file exists.

    python -m gpucsl_amd.tune --p 1000 --n 10000 --degree 4 --max-level 4 --out tuning.json
"""
from __future__ import annotations

import argparse
import ctypes as ct
import itertools
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


class LevelTuning(ct.Structure):
    _fields_ = [(f, ct.c_int32) for f in FIELDS]


def set_level(level: int, shape: dict) -> None:
    t = LevelTuning(*[int(shape[f]) for f in FIELDS])
    nat.call("gpucsl_set_level_tuning", level, ct.byref(t))


def get_level(level: int) -> dict:
    t = LevelTuning()
    nat.call("gpucsl_get_level_tuning", level, ct.byref(t))
    return {f: int(getattr(t, f)) for f in FIELDS}


def apply_table(table: dict) -> None:
    """Apply a tuning table; fields a table lacks keep their current values."""
    for lv, shape in table.get("levels", {}).items():
        cur = get_level(int(lv))
        cur.update({f: shape[f] for f in FIELDS if f in shape})
        set_level(int(lv), cur)


def _level_ms(timings: dict, level: int, reps: int) -> float:
    tot = 0.0
    for name in (f"gauss_level{level}_fast", f"gauss_level{level}_slow"):
        if name in timings:
            tot += timings[name][1]
    return tot / reps


def sweep(p: int, n: int, degree: float, max_level: int, alpha: float = 0.01, seed: int = 100, reps: int = 3,
          log=print) -> dict:
    import torch

    from gpucsl_amd import synthetic as syn
    from gpucsl_amd.gaussian import GaussianEngine
    from gpucsl_amd.pc.helpers import get_gaussian_thresholds

    data = syn.gaussian_sem(p, n, degree, seed=seed).data.astype(np.float32)
    eng = GaussianEngine(p, n, max_level, get_gaussian_thresholds(data, alpha))
    eng.set_data(torch.from_numpy(data).cuda())
    eng.run()
    torch.cuda.synchronize()
    want = eng.results()
    best = {lv: (float("inf"), None) for lv in range(1, max_level + 1)}
    combos = list(itertools.product(*[SPACE[f] for f in FIELDS]))
    t0 = time.perf_counter()
    for k, vals in enumerate(combos):
        shape = dict(zip(FIELDS, vals))
        set_level(0, shape)
        eng.run()
        torch.cuda.synchronize()
        with nat.KernelTimer() as kt:
            for _ in range(reps):
                eng.run()
            torch.cuda.synchronize()
        got = eng.results()
        if not ((got.skeleton == want.skeleton).all() and (got.separation_sets == want.separation_sets).all()):
            raise RuntimeError(f"shape {shape} changed the result")
        for lv in range(1, max_level + 1):
            ms = _level_ms(kt.timings, lv, reps)
            if ms < best[lv][0]:
                best[lv] = (ms, shape)
        if k % 100 == 0:
            log(f"[tune] {k + 1}/{len(combos)} shapes, {time.perf_counter() - t0:.1f}s")
    return {"workload": {"p": p, "n": n, "degree": degree, "max_level": max_level, "alpha": alpha, "seed": seed},
            "device": torch.cuda.get_device_name(0),
            "levels": {str(lv): dict(best[lv][1], ms=round(best[lv][0], 5)) for lv in best if best[lv][1]}}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--p", type=int, default=1000)
    ap.add_argument("--n", type=int, default=10000)
    ap.add_argument("--degree", type=float, default=4.0)
    ap.add_argument("--max-level", type=int, default=4)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--out", default="tuning.json")
    a = ap.parse_args(argv)
    nat.require_gpu()
    table = sweep(a.p, a.n, a.degree, a.max_level, reps=a.reps, log=lambda *x: print(*x, file=sys.stderr, flush=True))
    Path(a.out).write_text(json.dumps(table, indent=1))
    print(json.dumps(table))
    return 0


if __name__ == "__main__":
    sys.exit(main())
