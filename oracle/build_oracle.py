"""Builds the oracle's helpers (test infrastructure, never the product).

- oracle/_build/liboracle_rng.so  (g++, oracle/rng.cpp)
- oracle/_ref/pcaone_ref          (the reference's own path, oracle/ref.mk),
  only when the reference sources are present at REF (default /root/reference);
- oracle/_ref/pcaone_ref_amd      (the same path with libpcaone_amd bound in by
  pcaone_amd/binding/pcaone_reference_binding.hpp), when the library is built.
"""
from __future__ import annotations

import os
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent


def build_rng() -> Path:
    out = HERE / "_build" / "liboracle_rng.so"
    out.parent.mkdir(exist_ok=True)
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", str(HERE / "rng.cpp"), "-o", str(out)],
                   check=True)
    return out


def build_ref(ref: str | None = None, jobs: int = 8) -> Path | None:
    ref = ref or os.environ.get("PCAONE_REF", "/root/reference")
    if not (Path(ref) / "src" / "Halko.cpp").exists():
        return None
    subprocess.run(["make", "-s", "-f", "oracle/ref.mk", f"REF={ref}", f"-j{jobs}"], cwd=ROOT, check=True)
    if (ROOT / "pcaone_amd" / "libpcaone_amd.so").exists():
        # the reference path with the maintainer-side binding of libpcaone_amd
        subprocess.run(["make", "-s", "-f", "oracle/ref.mk", f"REF={ref}", f"-j{jobs}", "binding"], cwd=ROOT,
                       check=True)
    return HERE / "_ref" / "pcaone_ref"


if __name__ == "__main__":
    print(build_rng())
    print(build_ref())
