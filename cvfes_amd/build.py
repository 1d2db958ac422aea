"""Compile the gfx950 HIP sources into ``cvfes_amd/libcvfes_mi355x.so`` (in-tree).

Each ``csrc/*.hip`` file is compiled to an object with ``hipcc --offload-arch=gfx950`` (in
parallel), then linked into one shared library exporting the C ABI of
``include/cvfes_mi355x.h``.  Packed-FP32 instructions are disabled for all device code.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
ROOT = PKG_DIR.parent
CSRC = PKG_DIR / "csrc"
INCLUDE = ROOT / "include"
LIB_NAME = "libcvfes_mi355x.so"
LIB_PATH = PKG_DIR / LIB_NAME
ARCH = os.environ.get("CVFES_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _hipcc() -> str:
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(exe).exists():
        raise RuntimeError("hipcc not found: the HIP toolchain is required to build cvfes_amd")
    return exe


def _compile(src: Path, obj: Path) -> None:
    cmd = [_hipcc(), *HIP_FLAGS, f"-I{INCLUDE}", f"-I{CSRC}", "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{res.stdout}\n{res.stderr}")


def build(verbose: bool = False) -> Path:
    """Compile every HIP source and link ``libcvfes_mi355x.so``; returns its path."""
    objdir = PKG_DIR / "_build"
    objdir.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip"))
    objs = [objdir / (s.stem + ".o") for s in sources]
    workers = min(len(sources), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        list(ex.map(lambda so: _compile(*so), zip(sources, objs)))
    tmp = LIB_PATH.with_suffix(".so.tmp")
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB_PATH)
    if verbose:
        print(f"built {LIB_PATH}", file=sys.stderr)
    return LIB_PATH


if __name__ == "__main__":
    build(verbose=True)
