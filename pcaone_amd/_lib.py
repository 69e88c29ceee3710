"""ctypes binding of libpcaone_amd.so (the C ABI declared in include/pcaone_amd.h).

The library is loaded from the package directory only; a missing library is an
error, never a silent fallback.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libpcaone_amd.so"

PREC_FP64, PREC_FP32, PREC_BF16, PREC_BF16X1 = 0, 1, 2, 3
PRECISIONS = {"fp64": PREC_FP64, "fp32": PREC_FP32, "bf16": PREC_BF16, "bf16x1": PREC_BF16X1}


class PcaoneError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"[status {status}] {msg}")
        self.status = status


class SynthSpec(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("nsamples", C.c_uint32), ("npops", C.c_uint32), ("nsnps", C.c_uint64),
                ("freq_lo", C.c_double), ("freq_hi", C.c_double), ("fst_lo", C.c_double), ("fst_hi", C.c_double),
                ("miss_rate", C.c_double), ("high_miss_frac", C.c_double), ("high_miss_rate", C.c_double)]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is not built; run `python -m pcaone_amd.build` (or __graft_entry__.build())")
        L = C.CDLL(str(LIB_PATH))
        _declare(L)
        _lib = L
    return _lib


def _declare(L: C.CDLL) -> None:
    p = C.c_void_p
    L.pcaone_last_error.restype = C.c_char_p
    L.pcaone_last_error.argtypes = [p]
    L.pcaone_synth_packed.argtypes = [C.POINTER(SynthSpec), p, p]
    L.pcaone_synth_write_plink.argtypes = [C.POINTER(SynthSpec), C.c_char_p]
    for name in ("pcaone_synth_packed", "pcaone_synth_write_plink"):
        getattr(L, name).restype = C.c_int


def check(status: int, ctx=None) -> None:
    if status != 0:
        msg = lib().pcaone_last_error(ctx)
        raise PcaoneError(status, msg.decode() if msg else "unknown error")
