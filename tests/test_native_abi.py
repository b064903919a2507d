"""libgpucsl_hip.so loads and exports every symbol the header declares (CPU:
no compute calls)."""
import ctypes as ct
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "gpucsl_hip.h"


def declared():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[\w\s\*]*?\b(gpucsl_\w+)\s*\(", text, re.M)))


def test_header_parses():
    names = declared()
    assert "gpucsl_gauss_skeleton" in names and "gpucsl_last_error" in names
    assert len(names) >= 15


@pytest.mark.parametrize("which", ["LIB_PATH", "DEBUG_LIB_PATH"])
def test_library_exports_header_symbols(which):
    """The release library and its index-checked debug build export every
    entry point the header declares."""
    from gpucsl_amd import _native

    path = getattr(_native, which)
    if not path.exists():
        pytest.fail(f"{path.name} is not built (run __graft_entry__.build())")
    lib = ct.CDLL(str(path))
    for name in declared():
        assert hasattr(lib, name), name


def test_debug_library_selection_is_scoped():
    from gpucsl_amd import _native

    assert not _native._debug
    with _native.debug_library(True):
        assert _native._debug
        with _native.debug_library(False):
            assert _native._debug          # an inner non-debug call keeps the outer choice
    assert not _native._debug


def test_bindings_cover_header():
    from gpucsl_amd import _native

    assert set(_native.SIGNATURES) == set(declared())


def test_abi_version_without_gpu():
    from gpucsl_amd import _native

    lib = _native.load()
    assert lib.gpucsl_abi_version() == 2
    assert lib.gpucsl_max_built_level() == 32
    assert lib.gpucsl_max_built_level() >= 4


def test_product_does_not_import_oracle():
    for f in (ROOT / "gpucsl_amd").rglob("*.py"):
        src = f.read_text()
        assert "import oracle" not in src and "from oracle" not in src, f


def test_job_struct_layout_matches_header():
    """The ctypes mirror of gpucsl_gauss_job has the C layout: 4+4+8+4+4+4+4
    bytes of scalars, ten pointers, 33 doubles (x86-64 / gfx950 alignment)."""
    import ctypes as ct

    from gpucsl_amd.batch import GaussJob, MAX_LEVELS

    assert MAX_LEVELS == 33
    assert GaussJob.n.offset == 8 and GaussJob.flags.offset == 24 and GaussJob.X.offset == 32
    assert GaussJob.corr_ws.offset == 104 and GaussJob.thr.offset == 112 and ct.sizeof(GaussJob) == 112 + 33 * 8
    assert "#define GPUCSL_MAX_LEVELS 33" in (ROOT / "include" / "gpucsl_hip.h").read_text()
    src = (ROOT / "include" / "gpucsl_hip.h").read_text()
    body = src[src.index("typedef struct gpucsl_gauss_job"):src.index("} gpucsl_gauss_job;")]
    names = [ln.split(";")[0].replace("*", " ").split()[-1].split("[")[0]
             for ln in body.splitlines()[1:] if ";" in ln]
    assert names == [f[0] for f in GaussJob._fields_]


def test_disc_job_struct_layout_matches_header(tmp_path):
    """The ctypes mirror of gpucsl_disc_job has the field names of the header
    and the size and offsets the C compiler gives it; the batch size constant
    matches GPUCSL_DISC_BATCH_JOBS."""
    import ctypes as ct
    import subprocess

    from gpucsl_amd.discrete import BATCH_JOBS, DiscJob

    src = (ROOT / "include" / "gpucsl_hip.h").read_text()
    assert f"#define GPUCSL_DISC_BATCH_JOBS {BATCH_JOBS}" in src
    body = src[src.index("typedef struct gpucsl_disc_job"):src.index("} gpucsl_disc_job;")]
    names = [ln.split(";")[0].replace("*", " ").split()[-1] for ln in body.splitlines()[1:] if ";" in ln]
    assert names == [f[0] for f in DiscJob._fields_]
    prog = tmp_path / "layout.c"
    prog.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "gpucsl_hip.h"\nint main(void) {\n'
                    '  printf("%zu", sizeof(gpucsl_disc_job));\n'
                    + "".join(f'  printf(" %zu", offsetof(gpucsl_disc_job, {n}));\n' for n in names)
                    + "  return 0;\n}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), str(prog), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    assert got == [ct.sizeof(DiscJob)] + [getattr(DiscJob, n).offset for n in names]


def test_level_tuning_roundtrip_and_validation():
    """Launch shapes are host state of the library, per (bucket, level):
    settable for one bucket and level, every level or every bucket,
    range-checked, readable back (no GPU needed)."""
    from gpucsl_amd import _native as nat
    from gpucsl_amd.cli.cli_util import GpucslError
    from gpucsl_amd.tune import BUCKETS, get_shape, set_shape

    lib = nat.load()
    assert lib.gpucsl_tuning_bucket(1, 200) == BUCKETS["small"]
    assert lib.gpucsl_tuning_bucket(1, 1000) == BUCKETS["large"]
    assert lib.gpucsl_tuning_bucket(320, 500) == BUCKETS["batch"]
    assert lib.gpucsl_tuning_bucket(0, 10) == -1
    before = {(b, lv): get_shape(b, lv) for b in range(3) for lv in range(1, 9)}
    try:
        shape = dict(small_team_iters=2, wave_team_iters=4, fast_blocks_per_cu=8, fast_lds_bytes=16384,
                     item_tests=2048, wave_items=1)
        set_shape(1, 3, shape)
        assert get_shape(1, 3) == shape
        assert get_shape(0, 3) == before[(0, 3)] and get_shape(2, 3) == before[(2, 3)]
        set_shape(2, 0, dict(shape, fast_blocks_per_cu=2))
        assert all(get_shape(2, lv)["fast_blocks_per_cu"] == 2 for lv in range(1, 9))
        set_shape(-1, 4, dict(shape, item_tests=4096))
        assert all(get_shape(b, 4)["item_tests"] == 4096 for b in range(3))
        for bad in (dict(shape, fast_lds_bytes=70000), dict(shape, small_team_iters=0),
                    dict(shape, fast_lds_bytes=5000), dict(shape, item_tests=10), dict(shape, wave_items=2)):
            with pytest.raises(GpucslError):
                set_shape(1, 1, bad)
        with pytest.raises(GpucslError):
            set_shape(1, 9, shape)
        with pytest.raises(GpucslError):
            set_shape(3, 1, shape)
    finally:
        for (b, lv), sh in before.items():
            set_shape(b, lv, sh)


def test_tuning_table_buckets_apply():
    from gpucsl_amd import _native as nat
    from gpucsl_amd.tune import apply_table, get_shape, set_shape

    nat.load()
    before = {(b, lv): get_shape(b, lv) for b in range(3) for lv in range(1, 9)}
    try:
        apply_table({"buckets": {"batch": {"levels": {"2": {"fast_blocks_per_cu": 24}}},
                                 "small": {"levels": {"2": {"item_tests": 256}}}}})
        assert get_shape(2, 2)["fast_blocks_per_cu"] == 24 and get_shape(0, 2)["item_tests"] == 256
        assert get_shape(1, 2) == before[(1, 2)]
    finally:
        for (b, lv), sh in before.items():
            set_shape(b, lv, sh)


def test_corr_job_workspace_holds_digit_planes():
    """Host-only sizing: three 64-split column statistics arrays, the six int8
    digit planes (rows and columns padded to 64) and k_splits FP64 partial Gram
    matrices, each section 256-byte aligned."""
    from gpucsl_amd import _native

    lib = _native.load()
    al = lambda v: (v + 255) // 256 * 256
    for n, p, k in [(10000, 1000, 7), (333, 65, 1), (70000, 300, 2), (2, 1, 1)]:
        rows, cols = (n + 63) // 64 * 64, (p + 63) // 64 * 64
        want = al(3 * al(8 * 64 * p) + 6 * cols * rows) + 8 * k * p * p
        assert lib.gpucsl_corr_job_workspace_bytes(n, p, k) == al(want), (n, p, k)
    assert lib.gpucsl_corr_job_workspace_bytes(1, 10, 1) == 0      # n < 2
    assert lib.gpucsl_corr_job_workspace_bytes(100, 10, 0) == 0    # k_splits < 1
