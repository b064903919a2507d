"""ctypes bindings to libgpucsl_hip.so (the C ABI in include/gpucsl_hip.h).

There is no CPU fallback: if the shared library is missing or no GPU is
visible, every entry point raises GpucslError.  Buffers are torch tensors on
the current HIP device; the current torch stream is passed to every call.
"""
from __future__ import annotations

import ctypes as ct
from pathlib import Path

from gpucsl_amd.cli.cli_util import GpucslError

LIB_PATH = Path(__file__).resolve().parent / "libgpucsl_hip.so"
DEBUG_LIB_PATH = Path(__file__).resolve().parent / "libgpucsl_hip_debug.so"

_P = ct.c_void_p
_I = ct.c_int
_I64 = ct.c_int64
_SZ = ct.c_size_t
_D = ct.c_double
_F = ct.c_float

# name -> (restype, argtypes); must match include/gpucsl_hip.h
SIGNATURES = {
    "gpucsl_last_error": (ct.c_char_p, []),
    "gpucsl_abi_version": (_I, []),
    "gpucsl_max_built_level": (_I, []),
    "gpucsl_device_info": (_I, [_P, _P, _P]),
    "gpucsl_corr_workspace_bytes": (_SZ, [_I64, _I]),
    "gpucsl_corr": (_I, [_P, _I, _I64, _I, _P, _P, _P, _SZ, _P]),
    "gpucsl_corr_c64_to_c32": (_I, [_P, _I, _P, _P]),
    "gpucsl_gauss_workspace_bytes": (_SZ, [_I, _I]),
    "gpucsl_gauss_init": (_I, [_I, _I, _P, _P, _P, _P, _P]),
    "gpucsl_gauss_level0": (_I, [_P, _P, _I, _D, _F, _P, _P, _P, _P, _SZ, _P]),
    "gpucsl_gauss_job_workspace_bytes": (_SZ, [_I]),
    "gpucsl_corr_k_splits": (_I, [_I64, _I, _I]),
    "gpucsl_corr_job_workspace_bytes": (_SZ, [_I64, _I, _I]),
    "gpucsl_gauss_batch_workspace_bytes": (_SZ, [_P, _I]),
    "gpucsl_gauss_batch_corr": (_I, [_P, _P, _I, _P]),
    "gpucsl_gauss_batch_init": (_I, [_P, _P, _I, _P]),
    "gpucsl_gauss_batch_level0": (_I, [_P, _P, _I, _F, _P, _P]),
    "gpucsl_gauss_batch_level": (_I, [_I, _P, _P, _I, _F, _I, _P, _P, _SZ, _P]),
    "gpucsl_gauss_batch_finish": (_I, [_P, _P, _I, _P]),
    "gpucsl_gauss_batch_skeleton": (_I, [_P, _P, _I, _F, _I, _P, _P, _SZ, _P]),
    "gpucsl_gauss_level": (_I, [_I, _P, _P, _I, _I, _D, _F, _I, _P, _P, _P, _P, _P, _SZ, _P]),
    "gpucsl_gauss_skeleton": (_I, [_P, _P, _I, _I, _P, _F, _I, _P, _P, _P, _P, _P, _SZ, _P]),
    "gpucsl_gauss_skeleton_noalloc": (_I, [_P, _P, _I, _I, _P, _F, _I, _P, _P, _P, _P, _P]),
    "gpucsl_set_level_tuning": (_I, [_I, _P]),
    "gpucsl_get_level_tuning": (_I, [_I, _P]),
    "gpucsl_set_bucket_tuning": (_I, [_I, _I, _P]),
    "gpucsl_get_bucket_tuning": (_I, [_I, _I, _P]),
    "gpucsl_tuning_bucket": (_I, [_I, _I]),
    "gpucsl_pmax_from_zmin": (_I, [_P, _I, _I64, _P, _P]),
    "gpucsl_compact": (_I, [_P, _I, _P, _P, _SZ, _P]),
    "gpucsl_skeleton_to_u16": (_I, [_P, _I, _P, _P]),
    "gpucsl_gauss_outputs": (_I, [_P, _P, _I, _I64, _P, _P, _P]),
    "gpucsl_disc_encode": (_I, [_P, _I, _I64, _I, _I64, _I, _P, _I64, _P, _P, _P]),
    "gpucsl_unrank": (_I, [_I, _I, _P, _I, _P, _P]),
    "gpucsl_orient_edges": (_I, [_P, _I, _P, _I, _P, _P]),
    "gpucsl_orient_edges_arcs": (_I, [_P, _I, _P, _I, _P, _P, _I64, _P, _P]),
    "gpucsl_disc_workspace_bytes": (_SZ, [_I, _I, _I]),
    "gpucsl_disc_min_workspace_bytes": (_SZ, [_I, _I, _I]),
    "gpucsl_disc_init": (_I, [_I, _I, _I, _P, _P, _P, _P, _P, _SZ, _P]),
    "gpucsl_disc_level0": (_I, [_P, _I64, _I64, _P, _I, _I, _I, _D, _P, _P, _P, _P, _SZ, _P]),
    "gpucsl_disc_level": (_I, [_I, _P, _I64, _I64, _P, _I, _I, _I, _D, _P, _P, _P, _P, _P, _SZ, _P]),
    "gpucsl_disc_finish": (_I, [_P, _I, _P, _P, _P]),
    "gpucsl_disc_skeleton": (_I, [_P, _I64, _I64, _P, _I, _I, _I, _D, _P, _P, _P, _P, _P, _SZ, _P]),
    "gpucsl_disc_ci_test": (_I, [_P, _I64, _I64, _P, _I, _I, _P, _I, _I, _P, _P, _P]),
    "gpucsl_disc_batch_init": (_I, [_P, _I, _P]),
    "gpucsl_disc_batch_level0": (_I, [_P, _I, _P]),
    "gpucsl_disc_batch_level": (_I, [_I, _P, _I, _P]),
    "gpucsl_disc_batch_finish": (_I, [_P, _I, _P]),
    "gpucsl_disc_batch_skeleton": (_I, [_P, _I, _P]),
    "gpucsl_profile_begin": (_I, []),
    "gpucsl_profile_end": (_I, [ct.c_char_p, _SZ]),
}

STAT_SLOTS = 8
NOALLOC_SCRATCH_BYTES = 64 << 20   # GPUCSL_NOALLOC_SCRATCH_BYTES
STAT_NAMES = ("executed", "total", "rechecked", "flipped", "max_err_ratio", "max_degree", "useful")

_libs = {}
_debug = False


def load():
    """The library calls go to (each loaded once): the release build, or its
    index-checked debug build while a ``debug_library(True)`` block is active
    (process-wide).  Raises GpucslError if it is not built."""
    key = "debug" if _debug else "release"
    lib = _libs.get(key)
    if lib is None:
        path = DEBUG_LIB_PATH if _debug else LIB_PATH
        if not path.exists():
            raise GpucslError(
                f"{path.name} is not built; run `python -m gpucsl_amd.build` "
                "(hipcc --offload-arch=gfx950)")
        lib = ct.CDLL(str(path))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _libs[key] = lib
        _apply_tuning_table()
    return lib


class debug_library:
    """Context manager: calls inside go to the debug build when ``on``
    (gpucsl's is_debug, which compiles its kernels with device asserts,
    gpucsl/pc/kernel_management.py:32)."""

    def __init__(self, on: bool):
        self.on = bool(on)

    def __enter__(self):
        global _debug
        self.prev = _debug
        _debug = self.on or _debug
        return self

    def __exit__(self, *exc):
        global _debug
        _debug = self.prev
        return False


TUNING_TABLE = Path(__file__).resolve().parent / "tuning.json"


def _apply_tuning_table() -> None:
    """Per-level launch shapes measured by gpucsl_amd.tune, when present."""
    if TUNING_TABLE.exists():
        import json

        from gpucsl_amd.tune import apply_table

        apply_table(json.loads(TUNING_TABLE.read_text()))


def last_error() -> str:
    return load().gpucsl_last_error().decode(errors="replace")


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        raise GpucslError(f"{what}: {last_error()} (status {rc})" if what else last_error())


def call(name: str, *args):
    """Call an int-returning entry point and raise on a non-zero status."""
    rc = getattr(load(), name)(*args)
    check(rc, name)


def require_gpu():
    import torch

    if not torch.cuda.is_available():
        raise GpucslError("no HIP device is visible; gpucsl_amd runs only on a GPU")
    return torch


def to_host(*tensors) -> list:
    """Device tensors -> numpy arrays the caller owns.  Each is copied into a
    fresh pinned host tensor (one DMA at the link's rate; a pageable
    destination goes through the runtime's staging at a tenth of it), all on
    the current stream, with one synchronisation.  An array keeps its pinned
    tensor alive; freed ones return to torch's pinned-memory cache."""
    torch = require_gpu()
    outs = []
    for t in tensors:
        h = torch.empty(tuple(t.shape), dtype=t.dtype, pin_memory=True)
        h.copy_(t, non_blocking=True)
        outs.append(h)
    torch.cuda.current_stream().synchronize()
    return [h.numpy() for h in outs]


def stream_handle(stream=None) -> int:
    import torch

    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


def ptr(t) -> int | None:
    return None if t is None else t.data_ptr()


class KernelTimer:
    """Context manager: per-kernel launch durations from hipEvents recorded on
    the launch stream by the library.  ``timings`` maps kernel name to
    (launches, total_ms, min_ms, max_ms)."""

    def __enter__(self):
        call("gpucsl_profile_begin")
        self.timings = {}
        return self

    def __exit__(self, *exc):
        buf = ct.create_string_buffer(1 << 16)
        rc = load().gpucsl_profile_end(buf, len(buf))
        check(rc, "gpucsl_profile_end")
        for line in buf.value.decode().splitlines():
            name, cnt, tot, mn, mx = line.split("\t")
            self.timings[name] = (int(cnt), float(tot), float(mn), float(mx))
        return False


def decode_stats(host_stats, max_level: int):
    """uint64 [(max_level+1) * 8] -> list of per-level dicts."""
    import struct

    out = []
    for l in range(max_level + 1):
        row = [int(v) for v in host_stats[l * STAT_SLOTS:(l + 1) * STAT_SLOTS]]
        d = dict(zip(STAT_NAMES, row))
        d["max_err_ratio"] = struct.unpack("<f", struct.pack("<I", row[4] & 0xFFFFFFFF))[0]
        out.append(d)
    return out
