"""Build recipe for libgpucsl_hip.so (gfx950 only).

``python -m gpucsl_amd.build`` compiles every HIP source under ``csrc/`` into
one in-tree shared library next to this file.  Each source is compiled to an
object in parallel, then linked.
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


def build(verbose: bool = False, extra_flags=()) -> Path:
    OBJ.mkdir(exist_ok=True)
    cc = hipcc()
    headers = list(CSRC.glob("*.h")) + list((PKG.parent / "include").glob("*.h"))
    hdr_mtime = max((h.stat().st_mtime for h in headers), default=0)

    def compile_one(src: Path) -> Path:
        obj = OBJ / (src.stem + ".o")
        cmd = [cc, *HIPFLAGS, *extra_flags, "-c", str(src), "-o", str(obj)]
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
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, OUT)
    del hdr_mtime
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
