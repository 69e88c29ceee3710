"""Synthetic Balding-Nichols PLINK inputs (host generator of libpcaone_amd)."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib


@dataclass(frozen=True)
class Spec:
    seed: int
    nsamples: int
    nsnps: int
    npops: int = 3
    freq_lo: float = 0.05
    freq_hi: float = 0.5
    fst_lo: float = 0.1
    fst_hi: float = 0.1
    miss_rate: float = 0.0
    high_miss_frac: float = 0.0
    high_miss_rate: float = 0.0

    def c(self) -> _lib.SynthSpec:
        return _lib.SynthSpec(self.seed, self.nsamples, self.npops, self.nsnps, self.freq_lo, self.freq_hi,
                              self.fst_lo, self.fst_hi, self.miss_rate, self.high_miss_frac, self.high_miss_rate)

    @property
    def bytes_per_snp(self) -> int:
        return (self.nsamples + 3) // 4


def packed(spec: Spec, with_attempts: bool = False):
    """The .bed payload (nsnps x ceil(N/4) uint8, SNP-major, no magic)."""
    out = np.empty((spec.nsnps, spec.bytes_per_snp), dtype=np.uint8)
    att = np.empty(spec.nsnps, dtype=np.uint32) if with_attempts else None
    cs = spec.c()
    _lib.check(_lib.lib().pcaone_synth_packed(C.byref(cs), out.ctypes.data,
                                              att.ctypes.data if att is not None else None))
    return (out, att) if with_attempts else out


def write_plink(spec: Spec, prefix: str) -> None:
    cs = spec.c()
    _lib.check(_lib.lib().pcaone_synth_write_plink(C.byref(cs), str(prefix).encode()))


def cohorts(n: int = 256, seed0: int = 4000, n_range=(500, 20000), m_range=(10000, 200000), pops=(2, 6)):
    """Batched workload: per cohort c, N and M log-uniform in the ranges and
    2..6 populations, drawn from PCG64(seed0 + c); genotype seed seed0 + c."""
    out = []
    for c in range(n):
        g = np.random.Generator(np.random.PCG64(seed0 + c))
        ns = int(round(np.exp(g.uniform(np.log(n_range[0]), np.log(n_range[1])))))
        ms = int(round(np.exp(g.uniform(np.log(m_range[0]), np.log(m_range[1])))))
        npops = int(g.integers(pops[0], pops[1] + 1))
        out.append(Spec(seed=seed0 + c, nsamples=ns, nsnps=ms, npops=npops, freq_lo=0.05, freq_hi=0.5,
                        fst_lo=0.1, fst_hi=0.1))
    return out


# Named workloads (shapes and generator settings of the benchmark configurations).
def config(cfg: int, **over) -> Spec:
    base = {
        1: dict(seed=1001, nsamples=2000, nsnps=20000, npops=3, freq_lo=0.05, freq_hi=0.5, fst_lo=0.1, fst_hi=0.1),
        2: dict(seed=1002, nsamples=10000, nsnps=500000, npops=5, freq_lo=0.01, freq_hi=0.5, fst_lo=0.05,
                fst_hi=0.15),
        3: dict(seed=1003, nsamples=50000, nsnps=1000000, npops=8, freq_lo=0.05, freq_hi=0.5, fst_lo=0.1,
                fst_hi=0.1),
        5: dict(seed=1005, nsamples=20000, nsnps=300000, npops=3, freq_lo=0.05, freq_hi=0.5, fst_lo=0.1,
                fst_hi=0.1, miss_rate=0.10, high_miss_frac=0.02, high_miss_rate=0.30),
    }[cfg]
    base.update(over)
    return Spec(**base)
