"""Build recipe for libgpucsl_hip.so (gfx950 only).

``python -m gpucsl_amd.build`` compiles every HIP source under ``csrc/`` into
one in-tree shared library next to this file, and the same sources with
-DGPUCSL_DEBUG (index checks in the kernels) into libgpucsl_hip_debug.so,
which ``is_debug=True`` selects.  Each source is compiled to an object in
parallel, then linked.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OUT = PKG / "libgpucsl_hip.so"
OUT_DEBUG = PKG / "libgpucsl_hip_debug.so"
OBJ = PKG / "_build"
ARCH = os.environ.get("GPUCSL_OFFLOAD_ARCH", "gfx950")

# -packed-fp32-ops: keep the compiler from forming packed FP32 VALU ops.
HIPFLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unknown-pragmas",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def sources():
    return sorted(CSRC.glob("*.hip"))


def build(verbose: bool = False, extra_flags=(), debug: bool = False) -> Path:
    """Compile and link one library: the release build, or with debug the
    index-checked one."""
    obj_dir = OBJ / ("debug" if debug else "release")
    out = OUT_DEBUG if debug else OUT
    obj_dir.mkdir(parents=True, exist_ok=True)
    cc = hipcc()
    flags = [*HIPFLAGS, *(["-DGPUCSL_DEBUG"] if debug else []), *extra_flags]

    def compile_one(src: Path) -> Path:
        obj = obj_dir / (src.stem + ".o")
        cmd = [cc, *flags, "-c", str(src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed for {src.name}:\n{r.stderr}")
        if verbose and r.stderr.strip():
            print(r.stderr, file=sys.stderr)
        return obj

    srcs = sources()
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(compile_one, srcs))
    tmp = out.with_suffix(".so.tmp")
    cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, out)
    return out


def build_all(verbose: bool = False) -> list:
    return [build(verbose), build(verbose, debug=True)]


if __name__ == "__main__":
    print(build_all(verbose="-v" in sys.argv))
