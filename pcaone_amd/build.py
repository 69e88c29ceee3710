"""Build recipe for libpcaone_amd.so (HIP kernels for gfx950 + C++ host side).

Every source under pcaone_amd/csrc is compiled by hipcc into pcaone_amd/_build
and linked into pcaone_amd/libpcaone_amd.so.  `build()` always recompiles.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libpcaone_amd.so"
INCLUDE = PKG.parent / "include"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-O3", "-std=c++20", "-fPIC", "-Wall", "-Wno-unused-function", f"-I{INCLUDE}", f"-I{CSRC}"]
# Device code: gfx950 only; packed FP32 instructions are disabled.
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]
# Host code: OpenMP, and AVX2/FMA for the small L x L host steps (the same
# x86-64-v3 baseline the reference's own build targets); no FMA contraction, so
# the bit-exact host arithmetic (the Omega draw) stays the library's.
HOST = ["-fopenmp"]
HOST_ARCH = ["-Xarch_host", "-march=x86-64-v3", "-Xarch_host", "-ffp-contract=off"]


def _sources():
    return sorted(CSRC.glob("*.cpp")) + sorted(CSRC.glob("*.hip"))


def _compile(src: Path) -> Path:
    obj = BUILD / (src.name + ".o")
    # hipcc compiles .cpp as HIP too: every source gets the gfx950 device flags.
    cmd = [HIPCC] + COMMON + DEVICE + HOST_ARCH
    cmd += ["-x", "hip"] if src.suffix == ".hip" else HOST
    cmd += ["-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(line for line in (res.stdout + res.stderr).splitlines()
                    if "packed-fp32-ops" not in line and line.strip())
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {src.name}\n{' '.join(cmd)}\n{out}")
    if out:
        print(out, file=sys.stderr)
    return obj


def build(jobs: int | None = None) -> Path:
    if BUILD.exists():
        shutil.rmtree(BUILD)
    BUILD.mkdir(parents=True)
    srcs = _sources()
    jobs = jobs or min(8, os.cpu_count() or 4, len(srcs))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    # rocSOLVER (symmetric eigen-solver of the exact-PCA Gram matrix) and rocBLAS (its handle)
    cmd = [HIPCC, "-shared", "-fopenmp", f"--offload-arch={ARCH}", "-o", str(tmp)] + [str(o) for o in objs] + [
        "-L/opt/rocm/lib", "-lrocsolver", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/llvm/lib"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed\n{' '.join(cmd)}\n{res.stdout}{res.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build())
