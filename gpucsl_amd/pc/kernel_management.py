"""Kernel handles, drop-in for gpucsl/pc/kernel_management.py.

The reference compiles CUDA sources with NVRTC at run time (get_module,
:15-43) and launches them per level through Kernel.__call__ (:138-151).
Here every kernel is prebuilt for gfx950 in libgpucsl_hip.so, so a Kernel is
a handle that checks its level bound, plans memory and calls the library's
entry point; creating one compiles nothing.  The names, constructor
arguments and failure behaviour follow the reference:

- calling a CI kernel above its max_level raises AssertionError (:229-230,
  :289-290);
- a DiscreteKernel whose workspace cannot fit ``memory_restriction`` (or the
  device, when none is given) raises RuntimeError (:434-442), and
  ``max_memory_size`` reports what it will use.

GaussianPC/DiscretePC accept a ``Kernels`` object; the Gaussian one may carry
``group_size`` and ``margin`` for the CI test.
"""
from __future__ import annotations

from abc import ABC, abstractmethod
from typing import Tuple

from gpucsl_amd import _native as nat
from gpucsl_amd.gaussian import DEFAULT_GROUP_SIZE, DEFAULT_MARGIN


class Kernel(ABC):
    def __init__(self, is_debug: bool = False, should_log: bool = False):
        self.is_debug = is_debug
        self.should_log = should_log
        self._lib = nat.load()

    @abstractmethod
    def entry_point_for_level(self, level: int) -> str:
        """Name of the C-ABI function that runs this kernel at ``level``."""

    def pre_kernel_launch_check(self, level: int) -> None:
        pass

    def get_function_for_level(self, level: int):
        """The library function for ``level`` (ctypes), with its argtypes set."""
        return getattr(self._lib, self.entry_point_for_level(level))

    def launch(self, level: int, parameters: Tuple):
        fn = self.get_function_for_level(level)
        rc = fn(*parameters)
        nat.check(rc, self.entry_point_for_level(level))

    def __call__(self, level: int, *parameters):
        self.pre_kernel_launch_check(level)
        return self.launch(level, parameters)


class CompactKernel(Kernel):
    """Adjacency compaction (graph_helpers.cu:35-56 -> gpucsl_compact)."""

    def __init__(self, variable_count: int, device_count: int = 1, is_debug: bool = False,
                 should_log: bool = False):
        super().__init__(is_debug, should_log)
        self.variable_count = variable_count
        self.device_count = device_count

    def entry_point_for_level(self, level: int) -> str:
        return "gpucsl_compact"


class GaussianKernel(Kernel):
    """Fisher-z CI tests: level 0 -> gpucsl_gauss_level0, level l ->
    gpucsl_gauss_level (gaussian_ci.cu:18-120)."""

    def __init__(self, variable_count: int, device_count: int, max_level: int, is_debug: bool = False,
                 should_log: bool = False, group_size: int = DEFAULT_GROUP_SIZE, margin: float = DEFAULT_MARGIN):
        super().__init__(is_debug, should_log)
        self.variable_count = variable_count
        self.device_count = device_count
        self.max_level = max_level
        self.group_size = group_size
        self.margin = margin
        self.max_memory_size = int(self._lib.gpucsl_gauss_workspace_bytes(variable_count, max_level))

    def entry_point_for_level(self, level: int) -> str:
        return "gpucsl_gauss_level0" if level == 0 else "gpucsl_gauss_level"

    def get_function_for_level(self, level: int):
        self.pre_kernel_launch_check(level)
        return super().get_function_for_level(level)

    def pre_kernel_launch_check(self, level: int) -> None:
        assert level <= self.max_level, f"level {level} above max_level {self.max_level}"


class DiscreteKernel(Kernel):
    """Pearson chi^2 CI tests (discrete_ci.cu:37-246) with the workspace plan of
    find_execution_specification."""

    def __init__(self, max_level: int, variable_count: int, max_dim: int, n_observations: int,
                 memory_restriction: int = None, is_debug: bool = False, should_log: bool = False):
        super().__init__(is_debug, should_log)
        self.max_level = max_level
        self.variable_count = variable_count
        self.max_dim = max_dim
        self.n_observations = n_observations
        from gpucsl_amd.discrete import disc_workspace

        limit = memory_restriction if memory_restriction is not None else _device_memory()
        if limit is not None:
            limit = int(limit) - 256      # strictly below the restriction
        # RuntimeError when not even one test of max_level fits
        self.max_memory_size = disc_workspace(variable_count, max_level, max(max_dim, 1), limit)

    def entry_point_for_level(self, level: int) -> str:
        return "gpucsl_disc_level0" if level == 0 else "gpucsl_disc_level"

    def get_function_for_level(self, level: int):
        self.pre_kernel_launch_check(level)
        return super().get_function_for_level(level)

    def pre_kernel_launch_check(self, level: int) -> None:
        assert level <= self.max_level, f"level {level} above max_level {self.max_level}"


def _device_memory():
    """Total HBM of the current device, or None without a visible GPU."""
    try:
        import torch

        if torch.cuda.is_available():
            return int(torch.cuda.get_device_properties(torch.cuda.current_device()).total_memory)
    except Exception:
        pass
    return None


class Kernels:
    """gpucsl/pc/kernel_management.py:292-342."""

    def __init__(self):
        self.ci_test = None
        self.compact = None

    @staticmethod
    def for_gaussian_ci(variable_count: int, device_count: int, max_level: int, is_debug: bool = False,
                        should_log: bool = False):
        instance = Kernels()
        instance.ci_test = GaussianKernel(variable_count, device_count, max_level, is_debug, should_log)
        instance.compact = CompactKernel(variable_count, device_count, is_debug, should_log)
        return instance

    @staticmethod
    def for_discrete_ci(max_level: int, variable_count: int, max_dim: int, n_observations: int,
                        memory_restriction: int = None, is_debug: bool = False, should_log: bool = False):
        instance = Kernels()
        instance.ci_test = DiscreteKernel(max_level, variable_count, max_dim, n_observations, memory_restriction,
                                          is_debug, should_log)
        instance.compact = CompactKernel(variable_count, 1, is_debug, should_log)
        return instance
