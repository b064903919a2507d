"""Public PC API, drop-in for gpucsl/pc/pc.py: ``GaussianPC`` and ``DiscretePC``.

The class surface is gpucsl's (gpucsl/pc/pc.py:19-123): constructor
arguments, ``set_distribution_specific_options`` per test,
``discover_skeleton`` and ``execute`` returning ``TimedReturn(PCResult,
seconds)``.  A subclass only says which skeleton routine runs and with which
extra arguments; validation, the fully connected start graph and the
orientation step are shared.
"""
from abc import ABC, abstractmethod
from typing import List

import numpy as np

from gpucsl_amd.pc.edge_orientation.edge_orientation import orient_skeleton
from gpucsl_amd.pc.helpers import PCResult, init_pc_graph, timed
from gpucsl_amd.pc.pc_validations import (
    determine_max_level,
    validate_alpha,
    validate_given_devices,
    warn_against_too_high_memory_consumption_gaussian_pc,
)


class PC(ABC):
    def __init__(self, data: np.ndarray, max_level: int, alpha=0.05, kernels=None,
                 is_debug: bool = False, should_log: bool = False):
        validate_alpha(alpha)
        self.data = data
        self.num_observations, self.num_variables = data.shape[0], data.shape[1]
        self.max_level = determine_max_level(max_level, data)
        self.alpha = alpha
        self.kernels = kernels
        self.is_debug = is_debug
        self.should_log = should_log
        self.graph = init_pc_graph(data)

    def _common_arguments(self) -> dict:
        return dict(skeleton=self.graph, data=self.data, alpha=self.alpha, max_level=self.max_level,
                    num_variables=self.num_variables, num_observations=self.num_observations,
                    kernels=self.kernels, is_debug=self.is_debug, should_log=self.should_log)

    @abstractmethod
    def discover_skeleton(self):
        """TimedReturn(SkeletonResult, seconds) of this test's skeleton routine."""

    @abstractmethod
    def set_distribution_specific_options(self):
        """Options of one independence test; returns self."""

    @timed
    def execute(self):
        skeleton_run = self.discover_skeleton()
        sk = skeleton_run.result
        # the CPDAG of orient_edges(nx.DiGraph(skeleton), sepsets), from the
        # adjacency array directly
        orientation = orient_skeleton(sk.skeleton, sk.seperation_sets)
        return PCResult(orientation.result, sk.seperation_sets, sk.pmax, skeleton_run.runtime,
                        orientation.runtime, sk.discover_skeleton_kernel_time)


class GaussianPC(PC):
    """Fisher-z partial-correlation tests (gpucsl/pc/pc.py:71-102)."""

    def set_distribution_specific_options(self, devices: List[int] = [0], sync_device: int = None,
                                          correlation_matrix: np.ndarray = None, pcalg_compat: bool = False):
        """pcalg_compat (no counterpart in gpucsl): pcalg's thresholds, pMax and
        one-way separation sets instead of gpucsl's."""
        assert len(devices) > 0, "Invalid device list, must not be empty"
        warn_against_too_high_memory_consumption_gaussian_pc(self.max_level, self.num_variables)
        validate_given_devices(devices, sync_device)
        self.devices, self.sync_device = devices, sync_device
        self.correlation_matrix = correlation_matrix
        self.pcalg_compat = pcalg_compat
        return self

    def discover_skeleton(self):
        from gpucsl_amd.pc.discover_skeleton_gaussian import discover_skeleton_gpu_gaussian

        return discover_skeleton_gpu_gaussian(correlation_matrix=self.correlation_matrix, devices=self.devices,
                                              sync_device=self.sync_device, pcalg_compat=self.pcalg_compat,
                                              **self._common_arguments())


class DiscretePC(PC):
    """Pearson chi-square tests on categorical data (gpucsl/pc/pc.py:105-123)."""

    def set_distribution_specific_options(self, memory_restriction=None):
        self.memory_restriction = memory_restriction
        return self

    def discover_skeleton(self):
        from gpucsl_amd.pc.discover_skeleton_discrete import discover_skeleton_gpu_discrete

        return discover_skeleton_gpu_discrete(memory_restriction=self.memory_restriction,
                                              **self._common_arguments())
