"""Argument validation of the PC API (gpucsl/pc/pc_validations.py)."""
from __future__ import annotations

from gpucsl_amd.cli.cli_util import error, warning


def device_count() -> int:
    import torch

    return torch.cuda.device_count()


def validate_given_devices(devices, sync_device):
    """One GPU covers the whole adjacency matrix, so exactly one device id is
    accepted (the multi-device path of pc_validations.py:5-22 is not built)."""
    count = device_count()
    if len(devices) != 1:
        error(f"exactly one GPU is supported; got devices {list(devices)}")
    for d in devices:
        if d < 0 or d >= max(count, 1):
            error(f"Specified gpu indices should be between 0 and {max(count, 1) - 1}. You specified: {d}")
    if sync_device is not None and sync_device not in devices:
        error(f"The sync device has to be one of the specified gpus: {list(devices)}, got {sync_device}")


def validate_alpha(alpha):
    """pc_validations.py:25-27."""
    if alpha < 0 or alpha > 1:
        error("Alpha level has to be between 0 and 1")


def warn_against_too_high_memory_consumption_gaussian_pc(max_level: int, p: int | None = None):
    """Warn when the separation-set array alone would not fit the GPU
    (the reference's heuristic is in pc_validations.py:30-42)."""
    if p is None:
        return
    try:
        import torch

        total = torch.cuda.get_device_properties(torch.cuda.current_device()).total_memory
    except Exception:
        return
    if 4 * p * p * max(max_level, 1) > 0.5 * total:
        warning("the separation-set array needs more than half of the GPU memory; "
                "consider a smaller max level (-l/--level)")


def determine_max_level(max_level, data):
    """None -> p - 2; larger values are clamped to p - 2 (pc_validations.py:45-61)."""
    p = data.shape[1]
    max_possible = p - 2
    if max_level is None:
        max_level = max_possible
        if max_level < 0:
            error(f"Max level should be >= 0. Your input: {max_level}")
    if max_level > max_possible:
        warning(f"You set the max level to {max_level}, but the biggest possible level is "
                f"{max_possible}. The pc algorithm will at a maximum only run until level {max_possible}.")
        max_level = max_possible
    return max_level
