"""Builds libetm.so from the HIP sources in ``csrc/`` for gfx950, in-tree.

Every ``csrc/*.hip`` file is compiled to an object in ``_build/`` (in parallel), then linked into
``et_miner/lib/libetm.so``. The library ships with the source tree; nothing is installed.
"""

from __future__ import annotations

import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
ROOT = PKG_DIR.parent  # et-miner_amd/
CSRC = ROOT / "csrc"
INCLUDE = ROOT.parent / "include"
BUILD_DIR = ROOT / "_build"
LIB_DIR = PKG_DIR / "lib"
LIB_PATH = LIB_DIR / "libetm.so"

ARCH = "gfx950"
CXXFLAGS = [
    "-std=c++17",
    "-O3",
    f"--offload-arch={ARCH}",
    "-fPIC",
    # Keep the compiler from forming packed FP32 instructions (v_pk_*_f32).
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(path).exists():
        raise RuntimeError("hipcc not found: the ROCm toolchain is required to build libetm.so")
    return path


def _compile(src: Path, obj: Path, extra: list) -> None:
    cmd = [hipcc(), *CXXFLAGS, *extra, f"-I{INCLUDE}", f"-I{CSRC}", "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src.name}:\n{res.stderr[-6000:]}")


def build(verbose: bool = False, extra_flags: list | None = None, jobs: int | None = None) -> Path:
    """Compiles every HIP source and links libetm.so; returns its path. Always rebuilds."""
    extra = list(extra_flags or [])
    BUILD_DIR.mkdir(exist_ok=True)
    LIB_DIR.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip"))
    if not sources:
        raise RuntimeError(f"no HIP sources under {CSRC}")
    objs = [BUILD_DIR / (s.stem + ".o") for s in sources]
    workers = jobs or min(len(sources), max(1, min(8, os.cpu_count() or 1)))
    with ThreadPoolExecutor(workers) as pool:
        list(pool.map(lambda so: _compile(so[0], so[1], extra), zip(sources, objs)))
    tmp = LIB_PATH.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link of libetm.so failed:\n{res.stderr[-6000:]}")
    os.replace(tmp, LIB_PATH)
    if verbose:
        print(f"built {LIB_PATH} from {len(sources)} sources")
    return LIB_PATH


if __name__ == "__main__":
    build(verbose=True)
