"""Builds libxpbd_hip.so (all HIP kernels + host C++) in-tree for gfx950.

``python -m xpbd_amd.build`` compiles every source under ``xpbd_amd/csrc``
with hipcc (device code) or g++ (host-only C++) and links them into
``xpbd_amd/lib/libxpbd_hip.so``.  It always recompiles: it is the build
check, not an incremental build.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIBDIR = PKG / "lib"
LIB = LIBDIR / "libxpbd_hip.so"
OBJDIR = ROOT / "_build" / "obj"

ARCH = os.environ.get("XPBD_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: the kernels follow the reference's FP32 operation order,
# so no multiply-add may be fused behind their back.
# -munsafe-fp-atomics: float atomics use the hardware add instead of CAS loops.
# -packed-fp32-ops: packed FP32 (v_pk_*_f32) code generation is kept off.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-munsafe-fp-atomics", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
CXX_FLAGS = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off"]


def _flags_string() -> str:
    return " ".join(["-ffp-contract=off", "-munsafe-fp-atomics", "-packed-fp32-ops"])


def _compile(src: Path) -> Path:
    obj = OBJDIR / (src.name + ".o")
    inc = [f"-I{ROOT / 'include'}", f"-I{CSRC}", f'-DXPBD_BUILD_FLAGS="{_flags_string()}"']
    if src.suffix == ".hip":
        cmd = [HIPCC, *HIP_FLAGS, *inc, "-c", str(src), "-o", str(obj)]
    else:
        cmd = ["g++", *CXX_FLAGS, *inc, "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = True) -> Path:
    OBJDIR.mkdir(parents=True, exist_ok=True)
    LIBDIR.mkdir(parents=True, exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp"))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
