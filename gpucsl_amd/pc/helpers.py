"""Result types, timing and small host-side statistics.

Drop-in counterparts of gpucsl/pc/helpers.py: the same names, fields (including
the ``seperation_sets`` spelling) and tuple order, because callers unpack them.
"""
from __future__ import annotations

import logging
from functools import wraps
from math import sqrt
from timeit import default_timer as timer
from typing import Any, Callable, Dict, Generic, NamedTuple, Set, Tuple, TypeVar

import numpy as np

T = TypeVar("T")
U = TypeVar("U")


class TimedReturnT(Generic[T]):
    """Typing view of TimedReturn (gpucsl/pc/helpers.py:18-21)."""

    result: T
    runtime: float


class TimedReturn(NamedTuple):
    """(result, wall seconds) as returned by every @timed call (helpers.py:24-27)."""

    result: Any
    runtime: float


def timed(f: Callable[..., U]) -> Callable[..., TimedReturnT[U]]:
    """Wrap f so it returns TimedReturn(result, seconds) (helpers.py:30-39)."""

    @wraps(f)
    def wrapper(*args, **kwargs):
        t0 = timer()
        out = f(*args, **kwargs)
        return TimedReturn(out, timer() - t0)

    return wrapper


def get_gaussian_thresholds(data: np.ndarray, alpha: float = 0.05):
    """Per-level Fisher-z thresholds |Phi^-1(alpha/2)| / sqrt(n - l - 4).

    gpucsl/pc/helpers.py:42-50, including its off-by-one against pcalg's
    sqrt(n - l - 3): entry l uses n - (l + 1) - 3.  50 entries; entries for
    l >= min(50, n-3) - 2 stay 0.0, which never removes an edge.
    """
    from scipy.stats import norm

    n = data.shape[0]
    levels = 50
    thr = [0.0] * levels
    q = abs(norm.ppf(alpha / 2, loc=0, scale=1))
    for i in range(1, min(levels, n - 3) - 1):
        thr[i - 1] = q / sqrt(n - i - 3)
    return thr


def init_pc_graph(data: np.ndarray) -> np.ndarray:
    """Fully connected p x p uint16 start graph (helpers.py:53-55)."""
    p = data.shape[1]
    return np.ones((p, p), np.uint16)


def correlation_matrix_of(data: np.ndarray) -> np.ndarray:
    """Host FP64 Pearson correlation of the columns (helpers.py:58-59).  The
    skeleton functions compute it on the GPU instead; this stays for callers
    that precompute and pass ``correlation_matrix``."""
    return np.corrcoef(data, rowvar=False)


def transform_to_pmax(zmin: np.ndarray, num_samples: int) -> np.ndarray:
    """Host pmax = 2 (1 - Phi(sqrt(n-3) |zmin|)) (helpers.py:85-86)."""
    from scipy.stats import norm

    return 2 * norm.sf(np.abs(np.sqrt(num_samples - 3) * zmin))


def postprocess_pmax(pmax: np.ndarray) -> np.ndarray:
    """Set the diagonal and lower triangle to -1, as pcalg's pMax (helpers.py:96-101)."""
    pmax[np.tril_indices(pmax.shape[0])] = -1
    return pmax


def log(message):
    logging.info(message)


def log_time(message, value, unit="s", value_spacing=40):
    pad = value_spacing - len(message)
    suffix = ""
    if unit:
        pad -= len(unit) + 2
        suffix = f"({unit})"
    log(message + " " * max(pad, 0) + suffix + ": " + f"{value:1.6f}")


class PCResult(NamedTuple):
    """gpucsl/pc/helpers.py:138-144."""

    directed_graph: Any
    separation_sets: Dict[Tuple[int, int], Set[int]]
    pmax: np.ndarray
    discover_skeleton_runtime: float
    edge_orientation_runtime: float
    discover_skeleton_kernel_runtime: float


class SkeletonResult(NamedTuple):
    """gpucsl/pc/helpers.py:147-151."""

    skeleton: np.ndarray
    seperation_sets: Dict[Tuple[int, int], Set[int]]
    pmax: np.ndarray
    discover_skeleton_kernel_time: float
