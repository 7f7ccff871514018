"""Builds ``libedx.so`` (the HIP/CDNA4 C-ABI library) in-tree with hipcc.

Every ``*.hip`` file in this directory is compiled for gfx950 with hipcc,
every ``*.cpp`` file (host code only) with the host C++ compiler, and the
objects are linked into ``realspace_exatdiagonalization_py_amd/libedx.so``.  Object files go to
``csrc/_build/`` (git-ignored).  Run ``python -m
realspace_exatdiagonalization_py_amd.csrc.build`` or call ``build()``.
"""

from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

CSRC = Path(__file__).resolve().parent
PKG = CSRC.parent
ROOT = PKG.parent
LIB = PKG / "libedx.so"
ARCH = os.environ.get("EDX_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # element formulas keep the reference's separate roundings; fused ops are explicit fma()
    "-ffp-contract=off",
    # host dense algebra (std::complex) multiplies without the inf/NaN rescue call
    "-fcx-limited-range",
    # no packed FP32 instructions in generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    f"-I{ROOT / 'include'}",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found; the edx library needs ROCm's hipcc to build")


def _compile(src: Path, obj: Path) -> None:
    cmd = [_hipcc(), *COMMON, "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{r.stdout}\n{r.stderr}")


# host-only sources (no device code): the host C++ compiler
HOST_CXX = os.environ.get("EDX_HOST_CXX", "g++")
HOST_COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", f"-I{ROOT / 'include'}"]


def _compile_host(src: Path, obj: Path) -> None:
    cmd = [HOST_CXX, *HOST_COMMON, "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{HOST_CXX} failed for {src.name}:\n{r.stdout}\n{r.stderr}")


def build(verbose: bool = False) -> Path:
    """Compile all sources and link ``libedx.so``; returns its path."""
    out = CSRC / "_build"
    out.mkdir(exist_ok=True)
    jobs = [(_compile, s, out / (s.stem + ".o")) for s in sorted(CSRC.glob("*.hip"))]
    jobs += [(_compile_host, s, out / (s.stem + ".host.o")) for s in sorted(CSRC.glob("*.cpp"))]
    objs = [o for _, _, o in jobs]
    workers = max(1, min(len(jobs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4)), 16))
    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        list(ex.map(lambda j: j[0](j[1], j[2]), jobs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
