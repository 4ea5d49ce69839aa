"""Build of the engine's shared library for gfx950 with hipcc."""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PACKAGE_DIR = Path(__file__).resolve().parent
CSRC = PACKAGE_DIR / "csrc"
LIB_NAME = "libsquishy_volumes_amd.so"
LIB_PATH = PACKAGE_DIR / LIB_NAME
SOURCES = ["sort.hip", "transfer.hip", "closure.hip", "engine.hip", "frame_file.hip", "time_step.hip", "noalloc.hip", "collide.hip"]

# -packed-fp32-ops keeps the compiler from emitting packed FP32 instructions;
# -munsafe-fp-atomics lets float atomicAdd lower to the hardware atomics.
HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the engine is built with ROCm's hipcc for gfx950")


def build(verbose: bool = False, extra_flags: list[str] | None = None) -> Path:
    """Compile every HIP source into LIB_PATH (always recompiles)."""
    tmp = LIB_PATH.with_suffix(".so.tmp")
    cmd = [hipcc(), *HIPCC_FLAGS, *(extra_flags or []), "-o", str(tmp), *[str(CSRC / s) for s in SOURCES]]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed with exit code {res.returncode}")
    if verbose and res.stderr.strip():
        sys.stderr.write(res.stderr)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose=True))
