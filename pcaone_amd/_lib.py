"""ctypes binding of libpcaone_amd.so (the C ABI declared in include/pcaone_amd.h).

The library is loaded from the package directory only; a missing library is an
error, never a silent fallback.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libpcaone_amd.so"

PREC_FP64, PREC_FP32, PREC_BF16, PREC_BF16X1, PREC_BF16MIX, PREC_FP16, PREC_FP16MIX, PREC_INT8, PREC_FP8 = range(9)
PRECISIONS = {"fp64": PREC_FP64, "fp32": PREC_FP32, "bf16": PREC_BF16, "bf16x1": PREC_BF16X1,
              "bf16mix": PREC_BF16MIX, "fp16": PREC_FP16, "fp16mix": PREC_FP16MIX, "int8": PREC_INT8,
              "fp8": PREC_FP8}
ERR_NO_DEVICE = 6


class PcaoneError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"[status {status}] {msg}")
        self.status = status


class SynthSpec(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("nsamples", C.c_uint32), ("npops", C.c_uint32), ("nsnps", C.c_uint64),
                ("freq_lo", C.c_double), ("freq_hi", C.c_double), ("fst_lo", C.c_double), ("fst_hi", C.c_double),
                ("miss_rate", C.c_double), ("high_miss_frac", C.c_double), ("high_miss_rate", C.c_double)]


class Params(C.Structure):
    _fields_ = [("svd", C.c_int), ("k", C.c_uint32), ("oversamples", C.c_uint32), ("maxp", C.c_uint32),
                ("bands", C.c_uint32), ("maxiter", C.c_uint32), ("tol", C.c_double), ("tolem", C.c_double),
                ("maf", C.c_double), ("seed", C.c_int), ("rand", C.c_int), ("ploidy", C.c_int),
                ("noshuffle", C.c_int), ("emu", C.c_int), ("printv", C.c_int), ("verbose", C.c_int),
                ("precision", C.c_int), ("wm_g", C.c_int), ("wm_h", C.c_int), ("waves_g", C.c_int),
                ("waves_h", C.c_int), ("kchunk_g", C.c_uint64),
                ("kchunk_h", C.c_uint64), ("memory", C.c_double), ("buffer", C.c_uint32), ("ld", C.c_int),
                ("ld_stats", C.c_int), ("imaxiter", C.c_uint32), ("itol", C.c_double), ("stream", C.c_int),
                ("hardcall", C.c_int), ("threads", C.c_uint32), ("miss_gather", C.c_int)]


class RunInfo(C.Structure):
    _fields_ = [("epochs", C.c_int), ("total_epochs", C.c_int), ("em_iters", C.c_int), ("n_trace", C.c_int),
                ("n_sv", C.c_int), ("seconds", C.c_double), ("g_ms", C.c_double), ("h_ms", C.c_double),
                ("g_launches", C.c_uint64), ("h_launches", C.c_uint64),
                ("omega_ms", C.c_double), ("householder_fallbacks", C.c_uint64),
                ("gram_second_passes", C.c_uint64), ("epoch_rollbacks", C.c_uint64),
                ("stream_loads", C.c_uint64), ("stream_load_ms", C.c_double)]


_lib = None

_P = C.c_void_p
class BatchInfo(C.Structure):
    _fields_ = [("seconds", C.c_double), ("wm_g", C.c_int), ("waves_g", C.c_int), ("wm_h", C.c_int),
                ("waves_h", C.c_int), ("kchunk_g", C.c_uint64), ("kchunk_h", C.c_uint64),
                ("tuned_g_ms", C.c_double), ("tuned_h_ms", C.c_double), ("tuned_candidates", C.c_int),
                ("failed", C.c_int)]


_SIGS = {
    "pcaone_last_error": (C.c_char_p, [_P]),
    "pcaone_device_count": (C.c_int, [_P]),
    "pcaone_open": (C.c_int, [C.c_int, _P]),
    "pcaone_close": (C.c_int, [_P]),
    "pcaone_params_default": (None, [_P]),
    "pcaone_bed_read": (C.c_int, [_P, C.c_char_p, C.c_double, _P, _P]),
    "pcaone_bed_dims": (C.c_int, [C.c_char_p, _P, _P]),
    "pcaone_bed_reserve": (C.c_int, [_P, C.c_uint32, C.c_uint64]),
    "pcaone_bed_upload": (C.c_int, [_P, _P, C.c_uint64, C.c_uint64]),
    "pcaone_bed_finalize": (C.c_int, [_P]),
    "pcaone_bed_generate": (C.c_int, [_P, _P]),
    "pcaone_bed_download": (C.c_int, [_P, _P]),
    "pcaone_dims": (C.c_int, [_P, _P, _P, _P]),
    "pcaone_allele_freq": (C.c_int, [_P, _P]),
    "pcaone_standardize": (C.c_int, [_P, C.c_int]),
    "pcaone_permute_snps": (C.c_int, [_P, _P]),
    "pcaone_set_impute": (C.c_int, [_P, _P, _P, _P, C.c_uint32]),
    "pcaone_configure": (C.c_int, [_P, C.c_int, C.c_uint32]),
    "pcaone_set_omega": (C.c_int, [_P, _P]),
    "pcaone_sketch": (C.c_int, [_P, C.c_uint64, C.c_uint64, C.c_int]),
    "pcaone_update_omega": (C.c_int, [_P, C.c_int]),
    "pcaone_zero_h": (C.c_int, [_P, C.c_int]),
    "pcaone_finish_epoch": (C.c_int, [_P, C.c_uint32, C.c_int, _P, _P]),
    "pcaone_save_prev": (C.c_int, [_P]),
    "pcaone_get_u": (C.c_int, [_P, C.c_uint32, _P]),
    "pcaone_get_v": (C.c_int, [_P, C.c_uint32, _P]),
    "pcaone_run_halko": (C.c_int, [_P, _P, _P, _P, _P, _P, _P]),
    "pcaone_run_trace": (C.c_int, [_P, C.c_int, _P, C.c_int, _P]),
    "pcaone_batch_reserve": (C.c_int, [_P, C.c_uint32, _P, _P]),
    "pcaone_batch_upload": (C.c_int, [_P, C.c_uint32, _P, C.c_uint64, C.c_uint64]),
    "pcaone_batch_finalize": (C.c_int, [_P]),
    "pcaone_batch_generate": (C.c_int, [_P, C.c_uint32, _P]),
    "pcaone_batch_allele_freq": (C.c_int, [_P, C.c_uint32, _P]),
    "pcaone_batch_run": (C.c_int, [_P, _P, C.c_int, C.c_int, _P, _P, _P, _P, _P]),
    "pcaone_batch_get_u": (C.c_int, [_P, C.c_uint32, _P]),
    "pcaone_batch_get_v": (C.c_int, [_P, C.c_uint32, _P]),
    "pcaone_batch_trace": (C.c_int, [_P, C.c_uint32, _P, C.c_int, _P]),
    "pcaone_run_bfile": (C.c_int, [_P, C.c_char_p, C.c_char_p, _P, _P]),
    "pcaone_time_kernels": (C.c_int, [_P, C.c_int]),
    "pcaone_synth_packed": (C.c_int, [_P, _P, _P]),
    "pcaone_synth_write_plink": (C.c_int, [_P, C.c_char_p]),
    "pcaone_packed_gemm": (C.c_int, [_P, C.c_int, _P, C.c_uint64, C.c_uint32, _P, C.c_uint32, _P]),
    "pcaone_copy_probe": (C.c_int, [_P, C.c_uint64]),
    "pcaone_get_g": (C.c_int, [_P, C.c_uint64, C.c_uint64, _P]),
    "pcaone_get_h": (C.c_int, [_P, C.c_int, _P]),
    "pcaone_host_omega": (C.c_int, [C.c_int, C.c_int, C.c_uint64, C.c_uint64, _P]),
    "pcaone_thin_q_omega": (C.c_int, [_P, C.c_int]),
    "pcaone_block_plan": (C.c_int, [_P, C.c_uint32, C.c_uint64, _P, _P, _P, _P, C.c_uint64]),
    "pcaone_permute_plink_order": (C.c_int, [C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32, _P]),
    "pcaone_host_sym_eig": (C.c_int, [C.c_int, _P, _P, _P]),
    "pcaone_host_householder_signs": (C.c_int, [C.c_int, _P, _P]),
    "pcaone_host_chol_upper": (C.c_int, [C.c_int, _P, _P]),
    "pcaone_host_householder_q": (C.c_int, [C.c_uint64, C.c_int, _P]),
    "pcaone_residuals": (C.c_int, [_P, C.c_uint32, _P, _P, _P, C.c_int, C.c_uint64, C.c_uint64, _P]),
    "pcaone_stream_bed": (C.c_int, [_P, C.c_char_p, C.c_uint64, _P, _P]),
    "pcaone_stream_host": (C.c_int, [_P, _P, C.c_uint32, C.c_uint64, C.c_uint64]),
    "pcaone_packed_gemm_noalloc": (C.c_int, [_P, C.c_int, _P, C.c_uint64, C.c_uint32, _P, C.c_uint32, _P]),
    "pcaone_pgen_dims": (C.c_int, [C.c_char_p, _P, _P]),
    "pcaone_pgen_read": (C.c_int, [_P, C.c_char_p, C.c_double, C.c_int, _P, _P]),
    "pcaone_pgen_decode": (C.c_int, [C.c_char_p, C.c_int, C.c_uint64, C.c_uint64, _P]),
    "pcaone_stream_pgen": (C.c_int, [_P, C.c_char_p, C.c_int, C.c_uint64, _P, _P]),
    "pcaone_run_pgen": (C.c_int, [_P, C.c_char_p, C.c_char_p, _P, _P]),
    "pcaone_pgen_perm_order": (C.c_int, [C.c_uint64, C.c_uint32, C.c_uint64, C.c_uint32, C.c_int, _P]),
}


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is not built; run `python -m pcaone_amd.build` (or __graft_entry__.build())")
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status: int, ctx=None) -> None:
    if status != 0:
        msg = lib().pcaone_last_error(ctx)
        raise PcaoneError(status, msg.decode() if msg else "unknown error")
