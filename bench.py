"""Benchmark: Gaussian PC-stable skeleton, p=1000, n=10000 FP32, max level 4.

One step = correlation matrix of the resident data on MFMA + every level of
the skeleton + pmax, on one GPU, with inputs already in HBM.  Prints one JSON
line (see the keys below).  CI tests/s counts the tests a sequential
pcalg-order loop runs (device counter "useful"); the tests the GPU evaluated
(including lanes past the first pass of their iteration) are reported beside
it.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "CI tests/s + skeleton time, Gaussian p=1000 n=10k; p50/p99 request latency"
PEAK_HBM_GBS = 8000.0
PEAK_FP64_TFLOPS = 78.6
WORKLOADS = {
    # name: (p, n, degree, seed, alpha, max_level, dtype)
    "gaussian_p1000_n10000": (1000, 10000, 4.0, 100, 0.01, 4, np.float32),
    "gaussian_p50_n2000": (50, 2000, 3.0, 0, 0.01, 3, np.float64),
}
BATCH_WORKLOADS = ("batched_bootstrap256_p200_n5000_plus64",)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def level_bytes_per_test(level: int) -> int:
    """FP32 entries C_ij, C_ji, C[i,S], C[j,S], C[S,S] gathered by one level-l
    test (gpucsl/pc/cuda/helpers/gaussian_helpers.cu:56-76): 4 (l^2 + 2l + 2) B."""
    return 4 * (level * level + 2 * level + 2)


def load_traffic(kernel: str):
    f = ROOT / "profiles" / "traffic.json"
    if not f.exists():
        return None
    try:
        return json.loads(f.read_text()).get(kernel)
    except Exception:
        return None


def cpu_baseline(X: np.ndarray, n: int, alpha: float, max_level: int):
    """The FP64 oracle (C restatement of the reference's algorithm, OpenMP) on
    the host cores: correlation + skeleton of the same dataset."""
    import oracle

    oracle.lib()
    t0 = time.perf_counter()
    C = oracle.corrcoef(X.astype(np.float64))
    t1 = time.perf_counter()
    ref = oracle.gauss_skeleton(C, n, alpha, max_level)
    t2 = time.perf_counter()
    tests = int(ref["executed"].sum())
    return dict(value=tests / (t2 - t0), unit="CI tests/s", cores=oracle.num_threads(), kind="port",
                sample=(f"full workload once: corrcoef {t1 - t0:.3f}s + skeleton {t2 - t1:.3f}s, "
                        f"{tests} tests (FP64 oracle, OpenMP)"),
                seconds=t2 - t0, tests=tests)


def roofline(timings, profile_steps, stats, shapes, workload):
    """Roofline object for the kernel with the most time per step.  shapes:
    the (n, p) of every dataset one launch covers."""
    dominant = max(timings, key=lambda k: timings[k][1])
    cnt, tot, _, _ = timings[dominant]
    avg_s = tot / cnt / 1e3
    launches_per_step = cnt / profile_steps
    if dominant.startswith("corr_gram"):
        # symmetric rank-n update: n * p * (p + 1) flops for the p(p+1)/2 distinct entries
        flops = sum(1.0 * n * p * (p + 1) for n, p in shapes) / launches_per_step
        achieved = flops / avg_s / 1e12
        roof = dict(bound="mfma", achieved=achieved, peak=PEAK_FP64_TFLOPS, unit="TFLOP/s",
                    frac=achieved / PEAK_FP64_TFLOPS)
    elif dominant.startswith("gauss_level") and dominant.endswith("_fast"):
        lv = int(dominant[len("gauss_level")])
        # tests of this level per step / launches of this kernel per step = per launch
        bytes_per_launch = stats[lv]["executed"] * level_bytes_per_test(lv) / launches_per_step
        achieved = bytes_per_launch / avg_s / 1e9
        roof = dict(bound="hbm", achieved=achieved, peak=PEAK_HBM_GBS, unit="GB/s",
                    frac=achieved / PEAK_HBM_GBS)
    else:
        roof = dict(bound="hbm", achieved=None, peak=PEAK_HBM_GBS, unit="GB/s", frac=None)
    roof["kernel"] = dominant
    roof["avg_launch_ms"] = avg_s * 1e3
    # HBM bytes per launch from rocprofv3 FETCH_SIZE + WRITE_SIZE runs of this
    # workload (keys "kernel@workload"; bare kernel names are the default workload)
    roof["traffic"] = load_traffic(f"{dominant}@{workload}") or (
        load_traffic(dominant) if workload == "gaussian_p1000_n10000" else None)
    return roof


def batched_datasets(n_boot: int, n_sets: int):
    """Bootstrap resamples of one p=200, n=5000 dataset (degree 3, seed 300;
    resample b draws its rows with seed 300+b) and uneven datasets k with
    p ~ U[50, 500], n ~ U[1000, 20000] drawn with seed 400+k (degree 3, seed
    400+k).  Returns the base data, the row indices of every resample and the
    uneven datasets (float32)."""
    from gpucsl_amd import synthetic as syn

    base = syn.gaussian_sem(200, 5000, 3.0, seed=300).data
    rows = [np.random.Generator(np.random.PCG64(300 + b)).integers(0, 5000, size=5000) for b in range(n_boot)]
    sets = []
    for k in range(n_sets):
        rng = np.random.Generator(np.random.PCG64(400 + k))
        p, n = int(rng.integers(50, 501)), int(rng.integers(1000, 20001))
        sets.append(syn.gaussian_sem(p, n, 3.0, seed=400 + k).data.astype(np.float32))
    return base.astype(np.float32), rows, sets


def bench_batched(args):
    """Many datasets in shared launches (one captured graph per step) against
    the same datasets one after another through the single-dataset path."""
    import torch

    from gpucsl_amd import _native as nat
    from gpucsl_amd.batch import GaussJobState, GaussianBatch, run_plain

    alpha, L = 0.01, 3
    log(f"[bench] workload {args.workload}: generating {args.bootstraps} resamples + {args.datasets} datasets")
    base, rows, sets = batched_datasets(args.bootstraps, args.datasets)
    torch.cuda.set_device(0)
    base_dev = torch.from_numpy(base).cuda()
    jobs, shapes = [], []
    for r in rows:
        j = GaussJobState(5000, 200, L, alpha=alpha)
        j.set_data(base_dev[torch.from_numpy(r).cuda()])
        jobs.append(j)
        shapes.append((5000, 200))
    for X in sets:
        n, p = X.shape
        j = GaussJobState(n, p, L, alpha=alpha)
        j.set_data(X)
        jobs.append(j)
        shapes.append((n, p))
    torch.cuda.synchronize()
    hbm = sum(j.nbytes for j in jobs)
    log(f"[bench] {len(jobs)} jobs, {hbm / 2**30:.2f} GiB of job state")
    batch = GaussianBatch(jobs)
    batch.capture()

    log(f"[bench] batched: warmup {args.warmup}, timing {args.steps} graph replays")
    for _ in range(args.warmup):
        batch.run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.run()
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t0) / max(args.steps, 1)
    stats = batch.batch_stats().levels
    useful = sum(s["useful"] for s in stats)
    executed = sum(s["executed"] for s in stats)
    batched_results = [(j.view("skeleton", torch.int16, (j.p, j.p)).clone(),
                        j.view("sepsets", torch.int32, (j.p, j.p, max(j.max_level, 1))).clone()) for j in jobs]

    plain_steps = max(1, min(args.steps, 3))
    log(f"[bench] plain single-dataset path: {plain_steps} passes over the {len(jobs)} jobs")
    run_plain(jobs)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(plain_steps):
        run_plain(jobs)
    torch.cuda.synchronize()
    plain_sec = (time.perf_counter() - t0) / plain_steps
    same = all(torch.equal(a, j.view("skeleton", torch.int16, (j.p, j.p))) and
               torch.equal(b, j.view("sepsets", torch.int32, (j.p, j.p, max(j.max_level, 1))))
               for (a, b), j in zip(batched_results, jobs))

    log(f"[bench] profiled pass {args.profile_steps} eager batched steps")
    with nat.KernelTimer() as kt:
        for _ in range(args.profile_steps):
            batch.launch()
        torch.cuda.synchronize()
    timings = kt.timings
    roof = roofline(timings, args.profile_steps, stats, shapes, args.workload)

    cpu = None
    if not args.no_cpu_baseline:
        import oracle

        k = 4
        log(f"[bench] cpu baseline: oracle on {k} resamples")
        oracle.lib()
        t0 = time.perf_counter()
        tests = 0
        for r in rows[:k]:
            X = base[r].astype(np.float64)
            ref = oracle.gauss_skeleton(oracle.corrcoef(X), 5000, alpha, L)
            tests += int(ref["executed"].sum())
        cs = time.perf_counter() - t0
        cpu = dict(value=tests / cs, unit="CI tests/s", cores=oracle.num_threads(), kind="port",
                   sample=f"{k} bootstrap resamples (p=200, n=5000, L=3): corrcoef + skeleton, {tests} tests "
                          f"in {cs:.2f}s (FP64 oracle, OpenMP)")

    out = {
        "metric": METRIC,
        "value": useful / sec,
        "unit": "CI tests/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {args.bootstraps} bootstrap resamples of a p=200 n=5000 "
                               f"dataset + {args.datasets} datasets p~U[50,500] n~U[1000,20000], degree 3, "
                               f"float32 data, alpha={alpha}, max_level={L}; one captured graph per step",
                   "jobs": len(jobs), "job_state_gib": hbm / 2**30, "parallelism": "single GPU"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "plain_path": {"ms_per_pass": plain_sec * 1e3, "ci_tests_per_s": useful / plain_sec,
                       "batched_over_plain": plain_sec / sec, "results_identical": bool(same)},
        "tests_per_step": {"useful": useful, "executed": executed},
        "per_level": stats,
        "kernel_ms_per_step": {k: v[1] / args.profile_steps for k, v in timings.items()},
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="gaussian_p1000_n10000",
                    choices=sorted(WORKLOADS) + list(BATCH_WORKLOADS))
    ap.add_argument("--bootstraps", type=int, default=256, help="batched workload: bootstrap resamples")
    ap.add_argument("--datasets", type=int, default=64, help="batched workload: uneven datasets")
    ap.add_argument("--profile-steps", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="also run 3 device-to-device copies of MIB MiB (traffic-counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); the path shards across GPUs "
              "by design but this build runs on a single MI355X.", file=sys.stderr)
        sys.exit(2)

    if args.workload in BATCH_WORKLOADS:
        return bench_batched(args)

    import torch

    from gpucsl_amd import _native as nat
    from gpucsl_amd import synthetic as syn
    from gpucsl_amd.gaussian import GaussianEngine
    from gpucsl_amd.pc.helpers import get_gaussian_thresholds

    p, n, degree, seed, alpha, L, dtype = WORKLOADS[args.workload]
    log(f"[bench] workload {args.workload}: generating data")
    data = syn.gaussian_sem(p, n, degree, seed=seed).data.astype(dtype)
    torch.cuda.set_device(0)
    X = torch.from_numpy(np.ascontiguousarray(data)).cuda()
    thr = get_gaussian_thresholds(data, alpha)
    eng = GaussianEngine(p, n, L, thr)

    def step():
        eng.set_data(X)
        eng.run()

    if args.copy_calibration:
        nbytes = args.copy_calibration << 20
        src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"[bench] copy calibration: 3 copies of {args.copy_calibration} MiB")
        del src, dst

    log(f"[bench] warmup {args.warmup}")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    log(f"[bench] timing {args.steps} steps")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    sec = (t1 - t0) / max(args.steps, 1)
    stats = eng.level_stats()
    useful = sum(s["useful"] for s in stats)
    executed = sum(s["executed"] for s in stats)

    log(f"[bench] profiled pass {args.profile_steps} steps")
    with nat.KernelTimer() as kt:
        for _ in range(args.profile_steps):
            step()
        torch.cuda.synchronize()
    timings = kt.timings
    per_step_kernel_ms = {k: v[1] / args.profile_steps for k, v in timings.items()}
    roof = roofline(timings, args.profile_steps, stats, [(n, p)], args.workload)

    cpu = None
    if not args.no_cpu_baseline:
        log("[bench] cpu baseline (oracle)")
        cpu = cpu_baseline(data, n, alpha, L)

    out = {
        "metric": METRIC,
        "value": useful / sec,
        "unit": "CI tests/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: ER-DAG linear SEM degree {degree}, seed {seed}, "
                               f"{np.dtype(dtype).name} data, alpha={alpha}, max_level={L}",
                   "p": p, "n": n, "max_level": L, "alpha": alpha,
                   "correlation": "FP64 MFMA GEMM (C32 rounded from C64)",
                   "ci_tests": "FP32 with FP64 recheck near the threshold",
                   "parallelism": "single GPU"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "skeleton_ms": sec * 1e3,
        "tests_per_step": {"useful": useful, "executed": executed},
        "gpu_vs_cpu": (useful / sec) / cpu["value"] if cpu else None,
        "per_level": stats,
        "kernel_ms_per_step": per_step_kernel_ms,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
