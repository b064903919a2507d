"""Discrete skeleton discovery, drop-in for gpucsl/pc/discover_skeleton_discrete.py."""
from __future__ import annotations

from timeit import default_timer as timer

import numpy as np

from gpucsl_amd.cli.cli_util import GpucslError
from gpucsl_amd.pc.helpers import SkeletonResult, log, log_time, timed


@timed
def discover_skeleton_gpu_discrete(
    skeleton: np.ndarray,
    data: np.ndarray,
    alpha: float,
    max_level: int,
    num_variables: int,
    num_observations: int,
    kernels=None,
    memory_restriction: int = None,
    is_debug: bool = False,
    should_log: bool = False,
) -> SkeletonResult:
    """The skeleton routine above the C ABI (see the body below); with
    is_debug its calls go to the index-checked debug build of the library, as
    the reference compiles its kernels with device asserts
    (gpucsl/pc/kernel_management.py:32)."""
    from gpucsl_amd import _native as nat

    with nat.debug_library(is_debug):
        return _discrete_skeleton(skeleton=skeleton, data=data, alpha=alpha, max_level=max_level, num_variables=num_variables,
                                  num_observations=num_observations, kernels=kernels, memory_restriction=memory_restriction,
                                  is_debug=is_debug, should_log=should_log)


def _discrete_skeleton(
    skeleton: np.ndarray,
    data: np.ndarray,
    alpha: float,
    max_level: int,
    num_variables: int,
    num_observations: int,
    kernels=None,
    memory_restriction: int = None,
    is_debug: bool = False,
    should_log: bool = False,
) -> SkeletonResult:
    """Signature and result of gpucsl/pc/discover_skeleton_discrete.py:24-128.

    Numeric data are uploaded once and ordinal-encoded on the GPU into a padded
    column-major layout (other data are encoded on the host); every level runs
    on the GPU.  ``memory_restriction``
    bounds the workspace in bytes: below the full size fewer CI tests run at
    once, and a RuntimeError is raised if it cannot hold one test of max_level
    (the reference raises in the same situation,
    gpucsl/pc/kernel_management.py:434-442).
    """
    import torch

    from gpucsl_amd.discrete import DiscreteEngine, ordinal_encode, ordinal_encode_device

    if skeleton is not None and tuple(np.shape(skeleton)) != (num_variables, num_variables):
        raise GpucslError(f"skeleton must be {num_variables}x{num_variables}")
    dev_codes = ordinal_encode_device(data)
    if dev_codes is not None:
        codes_t, dims_t, dims = dev_codes
    else:
        codes, dims = ordinal_encode(data)
    max_dim = int(dims.max()) if dims.size else 1
    ci = getattr(kernels, "ci_test", None)
    if memory_restriction is None and ci is not None and getattr(ci, "max_memory_size", None):
        memory_restriction = ci.max_memory_size
    # a restriction below the minimum workspace raises RuntimeError; between the
    # minimum and the full size fewer tests run at once
    engine = DiscreteEngine(num_variables, num_observations, max_level, alpha, max_dim,
                            ws_limit=memory_restriction)
    if dev_codes is not None:
        engine.set_codes_device(codes_t, dims_t)
    else:
        engine.set_codes(codes, dims)
    torch.cuda.synchronize()
    start = timer()
    engine.run()
    res = engine.results()
    computation_time = timer() - start
    log_time("Computation time", computation_time)
    if should_log:
        for level, st in enumerate(res.stats):
            log(f"level {level}: {st}")
    return SkeletonResult(res.skeleton, res.separation_sets, res.pmax, computation_time)
