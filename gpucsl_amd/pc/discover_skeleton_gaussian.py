"""Gaussian skeleton discovery, drop-in for gpucsl/pc/discover_skeleton_gaussian.py.

The level loop runs inside libgpucsl_hip.so on one GPU: correlation on MFMA,
level 0, then compaction and the level-l CI kernels, with no host
synchronisation between levels.
"""
from __future__ import annotations

from timeit import default_timer as timer
from typing import List

import numpy as np

from gpucsl_amd.cli.cli_util import GpucslError
from gpucsl_amd.pc.helpers import SkeletonResult, get_gaussian_thresholds, log, log_time, timed
from gpucsl_amd.pc.pc_validations import validate_given_devices


def _tuning(kernels):
    """Group size and recheck margin carried by a Kernels object, if any."""
    from gpucsl_amd.gaussian import DEFAULT_GROUP_SIZE, DEFAULT_MARGIN

    k = kernels[0] if isinstance(kernels, (list, tuple)) and kernels else kernels
    ci = getattr(k, "ci_test", None)
    return (getattr(ci, "group_size", DEFAULT_GROUP_SIZE), getattr(ci, "margin", DEFAULT_MARGIN))


@timed
def discover_skeleton_gpu_gaussian(
    skeleton: np.ndarray,
    data: np.ndarray,
    correlation_matrix: np.ndarray,
    alpha: float,
    max_level: int,
    num_variables: int,
    num_observations: int,
    kernels=None,
    is_debug: bool = False,
    should_log: bool = False,
    devices: List[int] = [0],
    sync_device: int = None,
    pcalg_compat: bool = False,
) -> SkeletonResult:
    """The skeleton routine above the C ABI (see the body below); with
    is_debug its calls go to the index-checked debug build of the library, as
    the reference compiles its kernels with device asserts
    (gpucsl/pc/kernel_management.py:32)."""
    from gpucsl_amd import _native as nat

    with nat.debug_library(is_debug):
        return _gaussian_skeleton(skeleton=skeleton, data=data, correlation_matrix=correlation_matrix, alpha=alpha,
                                  max_level=max_level, num_variables=num_variables, num_observations=num_observations,
                                  kernels=kernels, is_debug=is_debug, should_log=should_log, devices=devices,
                                  sync_device=sync_device, pcalg_compat=pcalg_compat)


def _gaussian_skeleton(
    skeleton: np.ndarray,
    data: np.ndarray,
    correlation_matrix: np.ndarray,
    alpha: float,
    max_level: int,
    num_variables: int,
    num_observations: int,
    kernels=None,
    is_debug: bool = False,
    should_log: bool = False,
    devices: List[int] = [0],
    sync_device: int = None,
    pcalg_compat: bool = False,
) -> SkeletonResult:
    """Signature and result of gpucsl/pc/discover_skeleton_gaussian.py:116-172.

    ``skeleton`` contributes its shape only (level 0 sets every pair, as in
    gpucsl/pc/cuda/gaussian_ci.cu:39-50).  ``devices`` must name one GPU.
    ``pcalg_compat`` (no counterpart in gpucsl) switches to pcalg's
    conventions: thresholds |Phi^-1(alpha/2)| / sqrt(n - l - 3), pmax the largest
    p-value of any test (each with its own sqrt(n - |S| - 3)), and the sepset
    stored only for the ordered pair whose test removed the edge.
    """
    import torch

    from gpucsl_amd.gaussian import GaussianEngine

    validate_given_devices(devices, sync_device)
    if skeleton is not None and tuple(np.shape(skeleton)) != (num_variables, num_variables):
        raise GpucslError(f"skeleton must be {num_variables}x{num_variables}")
    if pcalg_compat:
        return _pcalg_skeleton(data, correlation_matrix, alpha, max_level, num_variables, num_observations,
                               devices[0], kernels)
    thresholds = get_gaussian_thresholds(data, alpha)
    group_size, margin = _tuning(kernels)
    with torch.cuda.device(devices[0]):
        start_init = timer()
        engine = GaussianEngine(num_variables, num_observations, max_level, thresholds,
                                margin=margin, group_size=group_size)
        if correlation_matrix is None:
            engine.set_data(data)
        else:
            engine.set_correlation(correlation_matrix)
        torch.cuda.synchronize()
        log_time("Initialization time", timer() - start_init)

        start = timer()
        if should_log:
            def on_level(level):
                torch.cuda.synchronize()
                log_time(f"    Level {level} time", timer() - start)

            engine.run_levels(on_level)
        else:
            engine.run()
        res = engine.results()
        computation_time = timer() - start
        log_time("Computation time", computation_time)
        if should_log:
            for level, st in enumerate(res.stats):
                log(f"level {level}: {st}")

    return SkeletonResult(res.skeleton, res.separation_sets, res.pmax, computation_time)


def _pcalg_skeleton(data, correlation_matrix, alpha, max_level, p, n, device, kernels) -> SkeletonResult:
    """pcalg conventions through the batched entry points (a batch of one job
    carrying GPUCSL_JOB_PCALG_COMPAT)."""
    import torch

    from gpucsl_amd.batch import GaussJobState, GaussianBatch

    group_size, margin = _tuning(kernels)
    with torch.cuda.device(device):
        x_dtype = "float64"
        if correlation_matrix is None:
            # the correlation is computed from the data as given: float32 data
            # on the float32 path, anything else in float64 (as the default mode)
            data = data if isinstance(data, torch.Tensor) else np.asarray(data)
            if data.dtype in (np.float32, torch.float32):
                x_dtype = "float32"
            elif not isinstance(data, torch.Tensor) and data.dtype != np.float64:
                data = data.astype(np.float64)
        job = GaussJobState(n, p, max_level, alpha=alpha, pcalg_compat=True, x_dtype=x_dtype)
        if correlation_matrix is None:
            job.set_data(data)
        else:
            job.set_correlation(correlation_matrix)
        batch = GaussianBatch([job], margin=margin, group_size=group_size)
        torch.cuda.synchronize()
        start = timer()
        batch.launch()
        res = job.results()
        computation_time = timer() - start
        log_time("Computation time", computation_time)
    return SkeletonResult(res.skeleton, res.separation_sets, res.pmax, computation_time)
