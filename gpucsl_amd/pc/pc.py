"""Public PC API, drop-in for gpucsl/pc/pc.py (GaussianPC, DiscretePC)."""
from abc import ABC, abstractmethod
from typing import List

import networkx as nx
import numpy as np

from gpucsl_amd.pc.edge_orientation.edge_orientation import orient_edges
from gpucsl_amd.pc.helpers import PCResult, init_pc_graph, timed
from gpucsl_amd.pc.pc_validations import (
    determine_max_level,
    validate_alpha,
    validate_given_devices,
    warn_against_too_high_memory_consumption_gaussian_pc,
)


class PC(ABC):
    """gpucsl/pc/pc.py:19-68."""

    def __init__(self, data: np.ndarray, max_level: int, alpha=0.05, kernels=None,
                 is_debug: bool = False, should_log: bool = False):
        validate_alpha(alpha)
        self.data = data
        self.max_level = determine_max_level(max_level, data)
        self.alpha = alpha
        self.kernels = kernels
        self.is_debug = is_debug
        self.should_log = should_log
        self.num_variables = data.shape[1]
        self.num_observations = data.shape[0]
        self.graph = init_pc_graph(data)

    @abstractmethod
    def discover_skeleton(self):
        pass

    @abstractmethod
    def set_distribution_specific_options(self):
        pass

    @timed
    def execute(self):
        (skeleton, separation_sets, pmax, kernel_time), skeleton_time = self.discover_skeleton()
        directed_graph, orientation_time = orient_edges(nx.DiGraph(skeleton), separation_sets)
        return PCResult(directed_graph, separation_sets, pmax, skeleton_time, orientation_time,
                        kernel_time)


class GaussianPC(PC):
    """gpucsl/pc/pc.py:71-102."""

    def discover_skeleton(self):
        from gpucsl_amd.pc.discover_skeleton_gaussian import discover_skeleton_gpu_gaussian

        return discover_skeleton_gpu_gaussian(
            self.graph, self.data, self.correlation_matrix, self.alpha, self.max_level,
            self.num_variables, self.num_observations, self.kernels, self.is_debug,
            self.should_log, self.devices, self.sync_device, self.pcalg_compat)

    def set_distribution_specific_options(self, devices: List[int] = [0], sync_device: int = None,
                                          correlation_matrix: np.ndarray = None, pcalg_compat: bool = False):
        """pcalg_compat (no counterpart in gpucsl): pcalg's thresholds, pMax and
        one-way separation sets instead of gpucsl's."""
        self.pcalg_compat = pcalg_compat
        warn_against_too_high_memory_consumption_gaussian_pc(self.max_level, self.num_variables)
        assert len(devices) > 0, "Invalid device list, must not be empty"
        validate_given_devices(devices, sync_device)
        self.devices = devices
        self.sync_device = sync_device
        self.correlation_matrix = correlation_matrix
        return self


class DiscretePC(PC):
    """gpucsl/pc/pc.py:105-123."""

    def discover_skeleton(self):
        from gpucsl_amd.pc.discover_skeleton_discrete import discover_skeleton_gpu_discrete

        return discover_skeleton_gpu_discrete(
            self.graph, self.data, self.alpha, self.max_level, self.num_variables,
            self.num_observations, self.kernels, self.memory_restriction, self.is_debug,
            self.should_log)

    def set_distribution_specific_options(self, memory_restriction=None):
        self.memory_restriction = memory_restriction
        return self
