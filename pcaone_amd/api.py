"""Python view of the C ABI: a device context holding one genotype matrix.

Every call goes to libpcaone_amd.so; there is no CPU fallback.  Arrays are
numpy, column-major where the C ABI is (U, V, Omega).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import PRECISIONS, Params, RunInfo, check, lib


def _ptr(a: np.ndarray):
    return a.ctypes.data if a is not None else None


def _f64(a) -> np.ndarray:
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


def default_params(**kw) -> Params:
    p = Params()
    lib().pcaone_params_default(C.byref(p))
    for k, v in kw.items():
        if k == "precision" and isinstance(v, str):
            v = PRECISIONS[v]
        if not hasattr(p, k):
            raise AttributeError(f"unknown parameter {k}")
        setattr(p, k, v)
    return p


def pgen_dims(prefix: str):
    n, m = C.c_uint32(), C.c_uint64()
    check(lib().pcaone_pgen_dims(str(prefix).encode(), C.byref(n), C.byref(m)))
    return n.value, m.value


def pgen_decode(prefix: str, j0: int = 0, j1: int | None = None, hardcall: bool = False) -> np.ndarray:
    """Variants [j0, j1) of <prefix>.pgen as .bed rows ((j1 - j0) x ceil(N/4)
    bytes), decoded on the host."""
    n, m = pgen_dims(prefix)
    j1 = m if j1 is None else j1
    out = np.empty((max(j1 - j0, 0), (n + 3) // 4), dtype=np.uint8)
    check(lib().pcaone_pgen_decode(str(prefix).encode(), int(hardcall), j0, j1, _ptr(out)))
    return out


def pgen_perm_order(nsnps: int, nbatches: int, blocksize: int, nthreads: int, seed: int = 112) -> np.ndarray:
    out = np.empty(max(nsnps, 1), dtype=np.int32)
    check(lib().pcaone_pgen_perm_order(nsnps, nbatches, blocksize, nthreads, seed, _ptr(out)))
    return out[:nsnps]


def device_count() -> int:
    n = C.c_int(0)
    check(lib().pcaone_device_count(C.byref(n)))
    return n.value


@dataclass
class Result:
    U: np.ndarray
    S: np.ndarray
    E: np.ndarray
    V: np.ndarray | None
    info: dict
    trace: list = field(default_factory=list)
    em_trace: list = field(default_factory=list)


class Context:
    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(lib().pcaone_open(device, C.byref(h)))
        self._h = h

    # -- lifecycle
    def close(self):
        if self._h:
            lib().pcaone_close(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st):
        check(st, self._h)

    # -- data
    def read_bed(self, prefix: str, maf: float = 0.0):
        n, m = C.c_uint32(), C.c_uint64()
        self._check(lib().pcaone_bed_read(self._h, str(prefix).encode(), maf, C.byref(n), C.byref(m)))
        return n.value, m.value

    def upload_packed(self, packed: np.ndarray, nsamples: int):
        packed = np.ascontiguousarray(packed, dtype=np.uint8)
        m = packed.shape[0]
        assert packed.shape[1] == (nsamples + 3) // 4
        self._check(lib().pcaone_bed_reserve(self._h, nsamples, m))
        self._check(lib().pcaone_bed_upload(self._h, _ptr(packed), 0, m))
        self._check(lib().pcaone_bed_finalize(self._h))

    def generate(self, spec):
        cs = spec.c()
        self._check(lib().pcaone_bed_generate(self._h, C.byref(cs)))

    def download_packed(self) -> np.ndarray:
        n, m, _ = self.dims()
        out = np.empty((m, (n + 3) // 4), dtype=np.uint8)
        self._check(lib().pcaone_bed_download(self._h, _ptr(out)))
        return out

    def dims(self):
        n, m, miss = C.c_uint32(), C.c_uint64(), C.c_uint64()
        self._check(lib().pcaone_dims(self._h, C.byref(n), C.byref(m), C.byref(miss)))
        return n.value, m.value, miss.value

    def allele_freq(self) -> np.ndarray:
        _, m, _ = self.dims()
        F = np.empty(m)
        self._check(lib().pcaone_allele_freq(self._h, _ptr(F)))
        return F

    def standardize(self, ploidy: int = 2):
        self._check(lib().pcaone_standardize(self._h, ploidy))

    def permute_snps(self, perm):
        p = np.ascontiguousarray(perm, dtype=np.int32)
        self._check(lib().pcaone_permute_snps(self._h, _ptr(p)))

    def set_impute(self, U, S, V):
        U, S, V = _f64(U), np.ascontiguousarray(S, dtype=np.float64), _f64(V)
        self._check(lib().pcaone_set_impute(self._h, _ptr(U), _ptr(S), _ptr(V), S.size))

    # -- operator steps
    def configure(self, precision, L: int):
        if isinstance(precision, str):
            precision = PRECISIONS[precision]
        self._check(lib().pcaone_configure(self._h, precision, L))

    def set_omega(self, omega: np.ndarray):
        o = _f64(omega)
        self._check(lib().pcaone_set_omega(self._h, _ptr(o)))

    def sketch(self, j0: int, j1: int, h_mode: int = 0):
        self._check(lib().pcaone_sketch(self._h, j0, j1, h_mode))

    def update_omega(self, source: int = 0):
        self._check(lib().pcaone_update_omega(self._h, source))

    def thin_q_omega(self, source: int = 0):
        self._check(lib().pcaone_thin_q_omega(self._h, source))

    def zero_h(self, which: int):
        self._check(lib().pcaone_zero_h(self._h, which))

    def finish_epoch(self, k: int, L: int, have_prev: bool = False):
        sig = np.empty(L)
        diff = C.c_double(0.0)
        self._check(lib().pcaone_finish_epoch(self._h, k, int(have_prev), _ptr(sig), C.byref(diff)))
        return sig, diff.value

    def save_prev(self):
        self._check(lib().pcaone_save_prev(self._h))

    def get_u(self, k: int) -> np.ndarray:
        n, _, _ = self.dims()
        U = np.empty((n, k), order="F")
        self._check(lib().pcaone_get_u(self._h, k, _ptr(U)))
        return U

    def get_v(self, k: int) -> np.ndarray:
        _, m, _ = self.dims()
        V = np.empty((m, k), order="F")
        self._check(lib().pcaone_get_v(self._h, k, _ptr(V)))
        return V

    def get_g(self, j0: int, j1: int, L: int) -> np.ndarray:
        G = np.empty((j1 - j0, L), order="F")
        self._check(lib().pcaone_get_g(self._h, j0, j1, _ptr(G)))
        return G

    def get_h(self, which: int, L: int) -> np.ndarray:
        n, _, _ = self.dims()
        H = np.empty((n, L), order="F")
        self._check(lib().pcaone_get_h(self._h, which, _ptr(H)))
        return H

    # ------------------------------------------------------------ batched form
    def batch_generate(self, specs):
        arr = (_lib.SynthSpec * len(specs))(*[sp.c() for sp in specs])
        self._check(lib().pcaone_batch_generate(self._h, len(specs), arr))
        self._batch_dims = [(sp.nsamples, sp.nsnps) for sp in specs]

    def batch_upload(self, cohorts):
        """cohorts: list of (packed rows (M x ceil(N/4) uint8), N)."""
        n = np.array([c[1] for c in cohorts], dtype=np.uint32)
        m = np.array([c[0].shape[0] for c in cohorts], dtype=np.uint64)
        self._check(lib().pcaone_batch_reserve(self._h, len(cohorts), _ptr(n), _ptr(m)))
        for i, (packed, nsamp) in enumerate(cohorts):
            packed = np.ascontiguousarray(packed, dtype=np.uint8)
            assert packed.shape[1] == (nsamp + 3) // 4
            self._check(lib().pcaone_batch_upload(self._h, i, _ptr(packed), 0, packed.shape[0]))
        self._check(lib().pcaone_batch_finalize(self._h))
        self._batch_dims = [(int(a), int(b)) for a, b in zip(n, m)]

    def batch_allele_freq(self, c: int) -> np.ndarray:
        F = np.empty(self._batch_dims[c][1])
        self._check(lib().pcaone_batch_allele_freq(self._h, c, _ptr(F)))
        return F

    def batch_run(self, params: Params | None = None, use_graph: bool = True, want_v: bool = False, **kw):
        p = params if params is not None else default_params(**kw)
        nc, k = len(self._batch_dims), p.k
        S = np.zeros((nc, k))
        E = np.zeros((nc, k))
        ep = np.zeros(nc, dtype=np.int32)
        stt = np.zeros(nc, dtype=np.int32)
        info = _lib.BatchInfo()
        self._check(lib().pcaone_batch_run(self._h, C.byref(p), int(use_graph), int(want_v), _ptr(S), _ptr(E),
                                           _ptr(ep), _ptr(stt), C.byref(info)))
        self._batch_k = k
        d = {f: getattr(info, f) for f, _ in _lib.BatchInfo._fields_}
        return {"S": S, "E": E, "epochs": ep, "status": stt, "info": d}

    def batch_get_u(self, c: int) -> np.ndarray:
        U = np.empty((self._batch_dims[c][0], self._batch_k), order="F")
        self._check(lib().pcaone_batch_get_u(self._h, c, _ptr(U)))
        return U

    def batch_get_v(self, c: int) -> np.ndarray:
        V = np.empty((self._batch_dims[c][1], self._batch_k), order="F")
        self._check(lib().pcaone_batch_get_v(self._h, c, _ptr(V)))
        return V

    def batch_trace(self, c: int):
        buf = np.empty(64)
        n = C.c_int(0)
        self._check(lib().pcaone_batch_trace(self._h, c, _ptr(buf), 64, C.byref(n)))
        return buf[: n.value].tolist()

    def copy_probe(self, nbytes: int):
        self._check(lib().pcaone_copy_probe(self._h, nbytes))

    def time_kernels(self, on: bool = True):
        self._check(lib().pcaone_time_kernels(self._h, int(on)))

    # -- drivers
    def run(self, params: Params | None = None, want_v: bool = False, **kw) -> Result:
        p = params if params is not None else default_params(**kw)
        n, m, _ = self.dims()
        k = p.k
        U = np.empty((n, k), order="F")
        S = np.zeros(k)
        E = np.zeros(k)
        V = np.empty((m, k), order="F") if want_v else None
        info = RunInfo()
        self._check(lib().pcaone_run_halko(self._h, C.byref(p), _ptr(U), _ptr(S), _ptr(E), _ptr(V), C.byref(info)))
        d = {f: getattr(info, f) for f, _ in RunInfo._fields_}
        return Result(U=U, S=S[: info.n_sv], E=E[: info.n_sv], V=V, info=d, trace=self._trace(0, info.n_trace),
                      em_trace=self._trace(1, max(info.em_iters, 0)))

    def _trace(self, em: int, n: int):
        buf = np.empty(max(n, 1))
        cnt = C.c_int(0)
        self._check(lib().pcaone_run_trace(self._h, em, _ptr(buf), n, C.byref(cnt)))
        return buf[: min(n, cnt.value)].tolist()

    def run_bfile(self, bfile: str, out: str, params: Params | None = None, **kw) -> dict:
        p = params if params is not None else default_params(**kw)
        info = RunInfo()
        self._check(lib().pcaone_run_bfile(self._h, str(bfile).encode(), str(out).encode(), C.byref(p),
                                           C.byref(info)))
        return {f: getattr(info, f) for f, _ in RunInfo._fields_}

    def run_pgen(self, pfile: str, out: str, params: Params | None = None, **kw) -> dict:
        """The --pgen path: read <pfile>.pgen/.pvar/.psam, run, write the files."""
        p = params if params is not None else default_params(**kw)
        info = RunInfo()
        self._check(lib().pcaone_run_pgen(self._h, str(pfile).encode(), str(out).encode(), C.byref(p),
                                          C.byref(info)))
        return {f: getattr(info, f) for f, _ in RunInfo._fields_}

    def read_pgen(self, prefix: str, maf: float = 0.0, hardcall: bool = False):
        n, m = C.c_uint32(), C.c_uint64()
        self._check(lib().pcaone_pgen_read(self._h, str(prefix).encode(), maf, int(hardcall), C.byref(n),
                                           C.byref(m)))
        return n.value, m.value

    def stream_pgen(self, prefix: str, block_snps: int, hardcall: bool = False):
        """Streamed genotypes from <prefix>.pgen (records decoded at every block load)."""
        n, m = C.c_uint32(), C.c_uint64()
        self._check(lib().pcaone_stream_pgen(self._h, str(prefix).encode(), int(hardcall), block_snps, C.byref(n),
                                             C.byref(m)))
        self._stream_rows = None
        return n.value, m.value

    def stream_bed(self, prefix: str, block_snps: int):
        """Streamed genotypes from <prefix>.bed (mapped; blocks loaded per product)."""
        n, m = C.c_uint32(), C.c_uint64()
        self._check(lib().pcaone_stream_bed(self._h, str(prefix).encode(), block_snps, C.byref(n), C.byref(m)))
        self._stream_rows = None
        return n.value, m.value

    def stream_host(self, packed: np.ndarray, nsamples: int, block_snps: int):
        """Streamed genotypes from host rows (M x ceil(N/4) bytes); the array is
        kept referenced by the context while it streams from it."""
        rows = np.ascontiguousarray(packed, dtype=np.uint8)
        self._check(lib().pcaone_stream_host(self._h, _ptr(rows), nsamples, rows.shape[0], block_snps))
        self._stream_rows = rows

    def residuals(self, U, S, V, j0: int = 0, j1: int | None = None, as_read: bool = False) -> np.ndarray:
        """Rows [j0, j1) of the centered residual matrix (float32, SNP-major);
        U = None subtracts nothing."""
        n, m, _ = self.dims()
        j1 = m if j1 is None else j1
        out = np.empty((j1 - j0, n), dtype=np.float32)
        if U is None:
            self._check(lib().pcaone_residuals(self._h, 0, None, None, None, int(as_read), j0, j1, _ptr(out)))
        else:
            U, S, V = _f64(U), np.ascontiguousarray(S, dtype=np.float64), _f64(V)
            self._check(lib().pcaone_residuals(self._h, U.shape[1], _ptr(U), _ptr(S), _ptr(V), int(as_read), j0, j1,
                                               _ptr(out)))
        return out

    def packed_gemm(self, precision, packed: np.ndarray, ncodes: int, W: np.ndarray) -> np.ndarray:
        if isinstance(precision, str):
            precision = PRECISIONS[precision]
        packed = np.ascontiguousarray(packed, dtype=np.uint8)
        W = _f64(W)
        Y = np.empty((packed.shape[0], W.shape[1]), order="F")
        self._check(lib().pcaone_packed_gemm(self._h, precision, _ptr(packed), packed.shape[0], ncodes, _ptr(W),
                                             W.shape[1], _ptr(Y)))
        return Y
