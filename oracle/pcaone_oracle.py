"""FP64 CPU oracle: a numpy restatement of PCAone's randomized-SVD path.

TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import this module, and only as the checker.  The
product (pcaone_amd) never calls it.

Every function cites the reference source it restates (PCAone v0.7.2, paths
relative to the reference root).  Linear algebra: numpy/LAPACK FP64.  LAPACK's
Householder QR (dgeqrf/dorgqr) uses the same reflector sign convention as
Eigen's HouseholderQR (beta = -sign(alpha)*||x||), so thin-Q factors match the
reference's up to rounding; the SVD of B may differ from Eigen's JacobiSVD by
column signs, which the reference's own outputs never fix on the plain path
(outputs are compared sign-aligned).

Pinning (see tests/test_oracle.py):
  decode / allele frequency / centering  - 256-byte KAT + brute force + golden .mbim
  standardize_E                          - closed form
  omega / shuffle                        - libstdc++ calls (oracle/rng.cpp); pinned by the
                                           --maxp 0 golden run (U after one pass from Omega)
  NormalOp / FancyOp / compute_usv       - golden outputs of the reference binary
                                           (tests/golden/*), dense SVD at convergence
  flip_omg / mev / flip_uv               - closed form + brute force
  fit_with_pi (EMU, incl. perm quirk)    - golden EMU outputs (-d 1 and -d 2)
  block_plan (out-of-core, -m)           - closed-form cases + golden -m runs (block count,
                                           band factor, trajectory)
  permute_plink_order                    - permutation invariants, the single-buffer closed form
                                           and golden -d 2 -m runs (loadings in file order)
  out-of-core NormalOp / FancyOp         - golden -m outputs of the reference binary
  pgen_perm_order (.pgen -m shuffle)     - permutation invariants, a hand-worked small case and
                                           the golden .pgen -d 2 -m runs of the reference binary
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
BED2GENO = np.array([1.0, -9.0, 0.5, 0.0])  # src/Common.hpp:58-59
VAR_TOL = 1e-9  # src/Data.hpp:7

_rng = None


def _rnglib():
    global _rng
    if _rng is None:
        path = HERE / "_build" / "liboracle_rng.so"
        if not path.exists():
            raise ImportError(f"{path} missing; run python oracle/build_oracle.py")
        _rng = C.CDLL(str(path))
        _rng.oracle_omega.argtypes = [C.c_int, C.c_int, C.c_long, C.c_long, C.c_void_p]
        _rng.oracle_shuffle.argtypes = [C.c_long, C.c_void_p]
        _rng.oracle_shuffle_parts.argtypes = [C.c_int, C.c_long, C.c_void_p, C.c_void_p]
    return _rng


def omega(seed: int, gaussian: bool, n: int, l: int) -> np.ndarray:
    """src/Halko.cpp:15-23, src/RSVD.hpp:25-59: N x L, column-major fill."""
    out = np.empty(n * l, dtype=np.float64)
    _rnglib().oracle_omega(int(seed), 1 if gaussian else 0, n, l, out.ctypes.data)
    return out.reshape(l, n).T.copy()


def shuffle_parts(seed: int, parts: list) -> list:
    """src/FilePgen.cpp:124-126: one std::default_random_engine seeded with
    `seed` shuffles each partition in turn (std::shuffle)."""
    sizes = np.array([len(p) for p in parts], dtype=np.int64)
    data = np.concatenate([np.asarray(p, dtype=np.int32) for p in parts]) if parts else np.zeros(0, np.int32)
    data = np.ascontiguousarray(data, dtype=np.int32)
    _rnglib().oracle_shuffle_parts(int(seed), len(parts), sizes.ctypes.data, data.ctypes.data)
    out, at = [], 0
    for s in sizes:
        out.append(data[at:at + s])
        at += s
    return out


def shuffle_indices(m: int) -> np.ndarray:
    """src/RSVD.hpp:61-70: permutation indices (permuted column i = original column idx[i])."""
    out = np.empty(m, dtype=np.int32)
    _rnglib().oracle_shuffle(m, out.ctypes.data)
    return out


# ---------------------------------------------------------------- PLINK decode
def read_plink(prefix: str):
    """src/FilePlink.hpp:10-30: N = lines(.fam), M = lines(.bim), magic 6c 1b 01."""
    prefix = str(prefix)
    with open(prefix + ".fam") as f:
        n = sum(1 for _ in f)
    with open(prefix + ".bim") as f:
        m = sum(1 for _ in f)
    raw = np.fromfile(prefix + ".bed", dtype=np.uint8)
    if raw.size < 3 or raw[0] != 0x6C or raw[1] != 0x1B or raw[2] != 0x01:
        raise ValueError("Incorrect magic number in plink bed file.")
    bps = (n + 3) >> 2
    packed = raw[3:3 + bps * m].reshape(m, bps)
    return packed, n, m


def decode_codes(packed: np.ndarray, n: int) -> np.ndarray:
    """2-bit codes, M x N (sample j at bits 2*(j%4) of byte j/4; src/FilePlink.cpp:74-83)."""
    shifts = np.array([0, 2, 4, 6], dtype=np.uint8)
    codes = (packed[:, :, None] >> shifts[None, None, :]) & 3
    return codes.reshape(packed.shape[0], -1)[:, :n]


def allele_freq(codes: np.ndarray) -> np.ndarray:
    """src/FilePlink.cpp:37-55: mean of BED2GENO over non-missing; 0 if none observed."""
    vals = BED2GENO[codes]
    obs = codes != 1
    cnt = obs.sum(axis=1)
    s = np.where(obs, vals, 0.0).sum(axis=1)
    return np.where(cnt > 0, s / np.maximum(cnt, 1), 0.0)


@dataclass
class Params:
    """Subset of src/Cmd.hpp Param used on the path, with src/Cmd.cpp:171-238 rules."""
    svd: int = 2
    k: int = 10
    oversamples: int = 10
    maxp: int = 20
    tol: float = 1e-4
    seed: int = 112
    rand: int = 1
    bands: int = 64
    noshuffle: bool = False
    emu: bool = False
    maxiter: int = 100
    tolem: float = 1e-5
    scale: int = -9
    ploidy: int = 2
    maf: float = 0.0
    memory: float = 0.0  # -m/--memory (GB): out-of-core when > 0
    buffer: int = 2  # --buffer (GB) of permute_plink
    ld: bool = False  # -D/--ld: write the residual matrix (src/Cmd.cpp:92)
    ld_stats: int = 0  # --ld-stats: 0 residuals of the PCs, 1 centered genotypes (src/Cmd.cpp:114)
    pgen: bool = False  # the input is a .pgen (src/Main.cpp:169-172: its -m shuffle is compute_pgen_perm)
    threads: int = 12  # -n/--threads (src/Cmd.hpp:31); the .pgen -m shuffle uses min(threads, cores)
    hardcall: bool = False  # --hardcall (src/Cmd.cpp:78)

    def __post_init__(self):
        self.oversamples = max(self.oversamples, self.k)  # src/Cmd.cpp:216
        self.out_of_core = self.memory > 0  # src/Cmd.cpp:218
        if self.maf > 0 and self.out_of_core:  # src/Cmd.cpp:221-224
            raise ValueError("does not support --maf filters for out-of-core mode yet! ")
        if not self.emu:
            self.maxiter = 0  # src/Cmd.cpp:230-232
        if self.bands < 4 or self.bands % 2 != 0:  # src/Cmd.cpp:235-236
            raise ValueError("the -w/--batches must be a power of 2 and the minimun is 4.")
        self.perm = self.svd == 2 and not self.noshuffle  # src/Cmd.cpp:238
        if self.svd == 3:
            self.out_of_core = False  # src/Cmd.cpp:218: -m does not apply to --svd 3
        self.missme = self.emu


@dataclass
class Data:
    """In-core Data for a PLINK input (src/Data.hpp, src/FilePlink.cpp:26-120)."""
    G: np.ndarray  # N x M centered (missing -> 0)
    F: np.ndarray
    C: np.ndarray | None  # M x N missing mask (missme only)
    params: Params
    perm: np.ndarray | None = None
    p_miss: float = 0.0
    keep: np.ndarray | None = None
    plan: "BlockPlan | None" = None
    G0: np.ndarray | None = None

    @property
    def nsamples(self):
        return self.G.shape[0]

    @property
    def nsnps(self):
        return self.G.shape[1]

    @classmethod
    def from_packed(cls, packed: np.ndarray, n: int, params: Params) -> "Data":
        codes = decode_codes(packed, n)
        order = None
        if params.out_of_core and params.perm:
            # src/Main.cpp:120-151: the file is rewritten in permute_plink order
            # before it is read, so F is estimated in that order; a .pgen is read
            # in compute_pgen_perm's logical order instead (src/Main.cpp:168-172,
            # F(snp_idx) of the logical index, src/FilePgen.cpp:229-234)
            if params.pgen:
                plan = block_plan(params, n, codes.shape[0])
                order = pgen_perm_order(codes.shape[0], params.bands, plan.blocksize,
                                        min(params.threads, os.cpu_count() or 1), params.seed)
            else:
                order = permute_plink_order(n, codes.shape[0], params.buffer, params.bands)
            codes = codes[order]
        F = allele_freq(codes)
        keep = None
        if params.maf > 0:  # src/Data.cpp:88-105 (MAF(a) = a > 0.5 ? 1-a : a, keep if MAF > maf)
            maf = np.where(F > 0.5, 1 - F, F)
            keep = np.nonzero(maf > params.maf)[0]
            codes = codes[keep]
            F = F[keep]
        vals = BED2GENO[codes]
        G = np.where(codes == 1, 0.0, vals - F[:, None]).T.copy()  # src/FilePlink.cpp:102-109
        C = (codes == 1) if (params.missme or params.out_of_core) else None
        p_miss = float(C.mean()) if C is not None else 0.0
        d = cls(G=G, F=F, C=C, params=params, p_miss=p_miss, keep=keep, perm=order)
        if params.out_of_core:
            d.plan = block_plan(params, n, codes.shape[0])
            d.G0 = G  # every block read starts from the centered codes
        return d

    def block_matrix(self, update, standardize, U, S, VT):
        """What read_block_initial / read_block_update (src/FilePlink.cpp:122-298)
        produce over all blocks: centered codes (missing -> 0); with update and
        --emu the missing entries take clamp(U S VT, -F, 1-F) with F of the SNP's
        own position; then the standardize scale sqrt(ploidy)/sd when sd > 1e-9."""
        A = self.G0.copy()
        if update and self.params.emu:
            pred = (U * S[None, :]) @ VT
            imp = np.minimum(np.maximum(pred, -self.F[None, :]), 1.0 - self.F[None, :])
            A = np.where(self.C.T, imp, A)
        if standardize and self.params.scale == -9:
            sd = np.sqrt(self.F * (1.0 - self.F))
            A = A * np.where(sd > VAR_TOL, math.sqrt(self.params.ploidy) / np.where(sd > VAR_TOL, sd, 1.0), 1.0)[None, :]
        return A

    def standardize_E(self):
        """src/Data.cpp:351-362."""
        if self.params.scale != -9:
            return
        sd = np.sqrt(self.F * (1.0 - self.F))
        fac = np.where(sd > VAR_TOL, math.sqrt(self.params.ploidy) / np.where(sd > VAR_TOL, sd, 1.0), 1.0)
        self.G *= fac[None, :]

    def fit_with_pi(self, U, S, VT):
        """src/Data.cpp:318-348 (EMU).  With the in-core shuffle the mask is read
        through perm (src/Data.cpp:322) while the clamp uses F(i) unpermuted (:328)."""
        if not self.params.emu:
            return
        pred = (U * S[None, :]) @ VT  # N x M
        if self.params.perm:
            mask = self.C[self.perm].T
        else:
            mask = self.C.T
        lo = -self.F[None, :]
        hi = 1.0 - self.F[None, :]
        imp = np.minimum(np.maximum(pred, lo), hi)  # fmin(fmax(x, -F), 1-F)
        self.G = np.where(mask, imp, self.G)


# ---------------------------------------------------------------- out-of-core
def _u32(x: int) -> int:
    return x & 0xFFFFFFFF


@dataclass
class BlockPlan:
    blocksize: int
    nblocks: int
    band_factor: int
    start: list
    stop: list


def block_plan(params: Params, nsamples: int, nsnps: int) -> BlockPlan:
    """Data::prepare's out-of-core plan (src/Data.cpp:42-84, Halko branch with
    dopca).  The RAM model 3nl + 2ml + 5m (doubles) is evaluated in 32-bit
    unsigned arithmetic, as the reference's uint counts (src/Data.hpp:43-45) do."""
    n, m = nsamples, nsnps
    l = params.k + params.oversamples
    t3, t2, t5 = _u32(3 * n * l), _u32(2 * m * l), _u32(5 * m)
    ram = float(_u32(t3 + t2 + t5)) / 134217728
    if params.memory > 1.1 * ram:
        ram = 0.0
    bs = math.ceil((((ram + params.memory) * 134217728 - t3) - t2 - t5) / n)
    if bs < 1:
        raise ValueError("out-of-core: -m/--memory leaves no room for a block of SNPs")
    nblocks = math.ceil(m / bs)
    if nblocks == 1:  # src/Data.cpp:63
        raise ValueError("only one block exists. please remove -m option")
    factor = 1
    if params.svd == 2:  # src/Data.cpp:64-76
        if nblocks < params.bands:
            bs = math.ceil(m / params.bands)
        else:
            factor = math.ceil(nblocks / params.bands)
            bs = math.ceil(m / (params.bands * factor))
        nblocks = math.ceil(m / bs)
    start = [i * bs for i in range(nblocks)]  # src/Data.cpp:79-83
    stop = [min(s + bs - 1, m - 1) for s in start]
    return BlockPlan(bs, nblocks, factor, start, stop)


def pgen_perm_order(nsnps: int, nbatches: int, blocksize: int, nthreads: int, seed: int) -> np.ndarray:
    """compute_pgen_perm (src/FilePgen.cpp:91-151): each batch of the SNP range
    is cut into blocks; a block's SNPs are split over the reader threads (the
    first len % active threads take one more); thread t reads a contiguous
    source partition of the file, shuffled (one engine seeded with --seed,
    partitions in thread order), and hands out its indices in turn."""
    if nsnps == 0:
        return np.zeros(0, dtype=np.int64)
    nbatches = max(1, min(nbatches, nsnps))
    blocksize = max(1, blocksize)
    nthreads = max(1, min(nthreads, min(blocksize, nsnps)))

    def blocks():
        for b in range(nbatches):
            bs, be = b * nsnps // nbatches, (b + 1) * nsnps // nbatches
            for s0 in range(bs, be, blocksize):
                ln = min(blocksize, be - s0)
                act = min(nthreads, ln)
                yield s0, [ln // act + (1 if t < ln % act else 0) for t in range(act)]

    count = [0] * nthreads
    for _, lens in blocks():
        for t, ln in enumerate(lens):
            count[t] += ln
    starts = np.cumsum([0] + count)
    parts = shuffle_parts(seed, [np.arange(starts[t], starts[t + 1]) for t in range(nthreads)])
    nxt = [0] * nthreads
    order = np.full(nsnps, -1, dtype=np.int64)
    for s0, lens in blocks():
        off = 0
        for t, ln in enumerate(lens):
            order[s0 + off:s0 + off + ln] = parts[t][nxt[t]:nxt[t] + ln]
            nxt[t] += ln
            off += ln
    return order


def permute_plink_order(nsamples: int, nsnps: int, gb: int, nbands: int) -> np.ndarray:
    """permute_plink (src/FilePlink.cpp:303-408) as an index map: position ib
    of the rewritten file holds original SNP order[ib].  Buffers of
    bufsize * nbands SNPs are read in turn; SNP j * nbands + b of buffer i is
    written at i * bufsize0 + bandidx[b] + j (bufsize0 = the first buffer's
    bufsize); the last, shorter buffer gives its remainder to the first bands."""
    bytes_per_snp = (nsamples + 3) >> 2
    two = int(math.floor(1073741824.0 * gb / bytes_per_snp))
    two = min(two, nsnps)
    bufsize = two // nbands
    if bufsize < 1:
        raise ValueError("permute_plink: the --buffer holds fewer SNPs than -w bands")
    two = bufsize * nbands
    nblocks = (nsnps + two - 1) // two
    modr2 = nsnps % two
    modr = nsnps % nbands
    bandsize = (nsnps + nbands - 1) // nbands
    bandidx = [i * bandsize if (modr == 0 or i < modr) else modr * bandsize + (bandsize - 1) * (i - modr)
               for i in range(nbands)]
    bufidx = bufsize
    order = np.full(nsnps, -1, dtype=np.int64)
    for i in range(nblocks):
        if i == nblocks - 1 and modr2 != 0:
            last = nsnps - (nblocks - 1) * two
            bufsize = (last + nbands - 1) // nbands
            modr2 = last % nbands
        for b in range(nbands):
            j = 0
            while j < bufsize - 1:
                order[i * bufidx + bandidx[b] + j] = i * two + j * nbands + b
                j += 1
            if i != nblocks - 1 or b < modr2 or modr2 == 0:
                order[i * bufidx + bandidx[b] + j] = i * two + j * nbands + b
    return order


# ---------------------------------------------------------------- helpers
def thin_q(X: np.ndarray) -> np.ndarray:
    """HouseholderQR(X).householderQ() * Identity(rows, cols) (src/Halko.cpp:121-122)."""
    return np.linalg.qr(X, mode="reduced")[0]


def full_piv_solve(A: np.ndarray, Bm: np.ndarray) -> np.ndarray:
    """X with A X = Bm by Eigen's FullPivHouseholderQR (the reference's
    `fullPivHouseholderQr().solve`, src/Halko.cpp:68): Householder QR with
    complete pivoting (external/Eigen/src/QR/FullPivHouseholderQR.h:470-530),
    early exit when the remaining corner is below eps * size of the first pivot,
    rank = pivots above eps * size * the largest |beta| (:246-255), and the
    basic solution: unknowns of the pivots past the rank are zero (:544-577).
    For an invertible A this is the ordinary solve."""
    qr = np.array(A, dtype=np.float64, copy=True)
    rows, cols = qr.shape
    size = min(rows, cols)
    prec = np.finfo(np.float64).eps * size
    rowt = np.arange(size)
    colt = np.arange(size)
    hc = np.zeros(size)
    nonzero, maxpivot, biggest = size, 0.0, 0.0
    for k in range(size):
        corner = np.abs(qr[k:, k:])
        r, c = np.unravel_index(np.argmax(corner), corner.shape)
        r += k
        c += k
        big = abs(qr[r, c])
        if k == 0:
            biggest = big
        if big <= biggest * prec:
            nonzero = k
            break
        rowt[k], colt[k] = r, c
        if r != k:
            qr[[k, r], k:] = qr[[r, k], k:]
        if c != k:
            qr[:, [k, c]] = qr[:, [c, k]]
        x = qr[k:, k]
        c0, tail = x[0], float(np.dot(x[1:], x[1:]))
        if tail <= np.finfo(np.float64).tiny:
            tau, beta = 0.0, c0
            x[1:] = 0.0
        else:
            beta = math.sqrt(c0 * c0 + tail)
            if c0 >= 0:
                beta = -beta
            x[1:] = x[1:] / (c0 - beta)
            tau = (beta - c0) / beta
        qr[k, k] = beta
        hc[k] = tau
        maxpivot = max(maxpivot, abs(beta))
        if k + 1 < cols:
            v = np.concatenate(([1.0], qr[k + 1:, k]))
            w = v @ qr[k:, k + 1:]
            qr[k:, k + 1:] -= tau * np.outer(v, w)
    perm = np.arange(cols)
    for k in range(size):
        perm[[k, colt[k]]] = perm[[colt[k], k]]
    diag = np.abs(np.diag(qr)[:nonzero])
    rank = int(np.sum(diag > maxpivot * prec))
    X = np.zeros((cols,) + Bm.shape[1:])
    if rank == 0:
        return X
    c = np.array(Bm, dtype=np.float64, copy=True)
    for k in range(rank):
        if rowt[k] != k:
            c[[k, rowt[k]]] = c[[rowt[k], k]]
        v = np.concatenate(([1.0], qr[k + 1:, k]))
        c[k:] -= hc[k] * np.outer(v, v @ c[k:])
    z = np.linalg.solve(np.triu(qr[:rank, :rank]), c[:rank])
    X[perm[:rank]] = z
    return X


def flip_omg(omg2: np.ndarray, omg: np.ndarray) -> np.ndarray:
    """src/RSVD.hpp:80-89; returns the new Omg (Omg2 := Omg is the caller's job)."""
    a = np.abs(omg2 - omg).sum(axis=0)
    b = np.abs(omg2 + omg).sum(axis=0)
    sgn = np.where(a > 2.0 * b, -1.0, 1.0)
    return omg * sgn[None, :]


def mev(X: np.ndarray, Y: np.ndarray) -> float:
    """src/Utils.cpp:194-200."""
    return float(np.linalg.norm(X.T @ Y, axis=0).sum() / X.shape[1])


def flip_uv(U: np.ndarray, V: np.ndarray, ubase: bool = False):
    """src/Utils.cpp:118-154 (column layout: U N x k, V M x k)."""
    U = U.copy()
    V = V.copy()
    base = U if ubase else V
    for i in range(base.shape[1]):
        x = int(np.argmax(np.abs(base[:, i])))  # first max, as Eigen maxCoeff
        if base[x, i] < 0:
            U[:, i] *= -1
            V[:, i] *= -1
    return U, V


# ---------------------------------------------------------------- operators
class RsvdOp:
    """RsvdOpData (src/Halko.hpp:6-42, src/Halko.cpp:15-97)."""

    def __init__(self, data: Data, k: int, os: int):
        self.data = data
        self.k = k
        self.os = os
        self.update = False
        self.standardize = False
        self.U = self.V = self.S = None
        self.trace: list[float] = []  # diff of every checked epoch
        self.epochs = 0
        self.init_omg()

    @property
    def L(self):
        return self.k + self.os

    def init_omg(self):
        p = self.data.params
        self.Omg = omega(p.seed, p.rand != 0, self.data.nsamples, self.L)
        self.Omg2 = self.Omg.copy()

    def set_flags(self, update, standardize):
        self.update, self.standardize = update, standardize

    def compute_gh(self, G, H, pi):
        raise NotImplementedError

    def compute_usv(self, p: int, tol: float):
        """src/Halko.cpp:46-97."""
        params = self.data.params
        k = self.k
        G = np.zeros((self.data.nsnps, self.L))
        H = np.zeros((self.data.nsamples, self.L))
        Upre = Ucur = None
        pi = 0
        while pi <= p:
            G, H = self.compute_gh(G, H, pi)
            Q1, R1 = np.linalg.qr(G, mode="reduced")
            Q2, R2 = np.linalg.qr(Q1, mode="reduced")
            R = R2 @ R1
            G = Q2
            B = full_piv_solve(R.T, H.T)
            UB, s, VBt = np.linalg.svd(B, full_matrices=False)
            Ucur = VBt.T[:, :k]
            self.epochs = pi + 1
            if pi > 0:
                diff = 1.0 - mev(Ucur, Upre)
                self.trace.append(diff)
                if diff < tol or pi == p:
                    if params.svd == 2 and 2.0 ** pi < params.bands:
                        p = int(math.log2(params.bands))
                    else:
                        self.U = Ucur
                        self.V = G @ UB[:, :k]
                        self.S = s[:k]
                        break
                else:
                    Upre = Ucur
            else:
                Upre = Ucur
            pi += 1
        self.U = Ucur


class NormalOp(RsvdOp):
    """NormalRsvdOpData::computeGandH in-core (src/Halko.cpp:99-127)."""

    def compute_gh(self, G, H, pi):
        d = self.data
        if d.params.out_of_core:
            return self._compute_gh_blocks(G, H, pi)
        if pi == 0:
            self.init_omg()
            if self.update:
                d.fit_with_pi(self.U, self.S, self.V.T)
            if self.standardize:
                d.standardize_E()
        if pi > 0:
            self.Omg = flip_omg(self.Omg2, thin_q(H))
            self.Omg2 = self.Omg.copy()
        G = d.G.T @ self.Omg
        H = d.G @ G
        return G, H

    def _compute_gh_blocks(self, G, H, pi):
        """src/Halko.cpp:129-152: no flipOmg; H summed over the blocks in order."""
        d = self.data
        if pi == 0:
            self.init_omg()
            self.A = d.block_matrix(self.update, self.standardize, self.U, self.S,
                                    None if self.V is None else self.V.T)
        if pi > 0:
            self.Omg = thin_q(H)
        G = np.zeros((d.nsnps, self.L))
        H = np.zeros((d.nsamples, self.L))
        for s, e in zip(d.plan.start, d.plan.stop):
            blk = slice(s, e + 1)
            G[blk] = self.A[:, blk].T @ self.Omg
            H += self.A[:, blk] @ G[blk]
        return G, H


class FancyOp(RsvdOp):
    """FancyRsvdOpData::computeGandH in-core (src/Halko.cpp:155-224).  Windows that
    start past the last SNP (possible when M < bands^2) are empty here; the
    reference computes a negative block size for them."""

    def __init__(self, data, k, os):
        super().__init__(data, k, os)
        self.H1 = np.zeros((data.nsamples, self.L))
        self.H2 = np.zeros((data.nsamples, self.L))
        self.bandsize = 1
        self.blocksize = 0
        self.omega_updates = 0

    def compute_gh(self, G, H, pi):
        d = self.data
        bands = d.params.bands
        if pi == 0:
            self.init_omg()
        if 2.0 ** pi >= bands:
            self.H1[:] = 0
            self.H2[:] = 0
        if d.params.out_of_core:
            return self._compute_gh_blocks(G, H, pi)
        if pi == 0:
            if self.update:
                d.fit_with_pi(self.U, self.S, self.V.T)
            if self.standardize:
                d.standardize_E()
            self.bandsize = 1
            self.blocksize = int(math.ceil(d.nsnps / bands))
            if d.params.perm and not (self.update and d.params.emu):
                d.perm = shuffle_indices(d.nsnps)
                d.G = d.G[:, d.perm].copy()
        self.bandsize = min(self.bandsize * 2, bands)
        bs = self.bandsize
        i = 1
        for b in range(bands):
            start = b * self.blocksize
            stop = d.nsnps - 1 if (b + 1) * self.blocksize >= d.nsnps else (b + 1) * self.blocksize - 1
            if start <= stop:
                blk = slice(start, stop + 1)
                G[blk] = d.G[:, blk].T @ self.Omg
                if i <= bs // 2:
                    self.H1 += d.G[:, blk] @ G[blk]
                else:
                    self.H2 += d.G[:, blk] @ G[blk]
            adjacent = pi > 0 and (b + 1) == 2 ** (pi - 1) and 2.0 ** pi < bands
            if (b + 1) < bs and not adjacent:
                i += 1
                continue
            if not (i == bs or i == bs // 2 or adjacent):
                i += 1
                continue
            H = self.H1 + self.H2
            self.Omg = flip_omg(self.Omg2, thin_q(H))
            self.Omg2 = self.Omg.copy()
            self.omega_updates += 1
            if i == bs:
                self.H1[:] = 0
                i = 0
            else:
                self.H2[:] = 0
            i += 1
        return G, H


    def _compute_gh_blocks(self, G, H, pi):
        """src/Halko.cpp:225-268: the plan's blocks are the windows, bands grow
        from band_factor blocks, flipOmg after every update."""
        d = self.data
        plan = d.plan
        bands = d.params.bands
        if pi == 0:
            self.A = d.block_matrix(self.update, self.standardize, self.U, self.S,
                                    None if self.V is None else self.V.T)
            self.bandsize = plan.band_factor
        self.bandsize = min(self.bandsize * 2, plan.nblocks)
        bs = self.bandsize
        i = 1
        for b in range(plan.nblocks):
            blk = slice(plan.start[b], plan.stop[b] + 1)
            G[blk] = self.A[:, blk].T @ self.Omg
            if i <= bs // 2:
                self.H1 += self.A[:, blk] @ G[blk]
            else:
                self.H2 += self.A[:, blk] @ G[blk]
            adjacent = pi > 0 and (b + 1) == 2 ** (pi - 1) * plan.band_factor and 2.0 ** pi < bands
            if (b + 1) < bs and not adjacent:
                i += 1
                continue
            if not (i == bs or i == bs // 2 or adjacent):
                i += 1
                continue
            H = self.H1 + self.H2
            self.Omg = flip_omg(self.Omg2, thin_q(H))
            self.Omg2 = self.Omg.copy()
            self.omega_updates += 1
            if i == bs:
                self.H1[:] = 0
                i = 0
            else:
                self.H2[:] = 0
            i += 1
        return G, H


@dataclass
class Result:
    U: np.ndarray
    S: np.ndarray
    V: np.ndarray
    E: np.ndarray
    epochs: int
    trace: list = field(default_factory=list)
    em_trace: list = field(default_factory=list)
    perm: np.ndarray | None = None
    residuals: np.ndarray | None = None  # M x N float32 in file (original SNP) order, with --ld


def run_pca_with_halko(data: Data) -> Result:
    """src/Halko.cpp:271-345 (in-core, genetic input, no LD)."""
    p = data.params
    op = FancyOp(data, p.k, p.oversamples) if p.svd == 2 else NormalOp(data, p.k, p.oversamples)
    em_trace = []
    if not p.missme:
        op.set_flags(False, not p.ld)  # src/Halko.cpp:283-285: no standardization with --ld
        op.compute_usv(p.maxp, p.tol)
    else:
        op.set_flags(False, False)
        op.compute_usv(p.maxp, p.tol)
        op.U, op.V = flip_uv(op.U, op.V, False)
        for _ in range(p.maxiter):
            op.set_flags(True, False)
            Vpre = op.V
            op.compute_usv(p.maxp, p.tol)
            op.U, op.V = flip_uv(op.U, op.V, False)
            diff = 1.0 - mev(op.V, Vpre)
            em_trace.append(diff)
            if diff < p.tolem:
                break
        if p.emu:
            op.set_flags(True, True)
            op.compute_usv(p.maxp, p.tol)
            op.U, op.V = flip_uv(op.U, op.V, False)
    S = op.S if op.S is not None else np.zeros(0)
    res = Result(U=op.U, S=S, V=op.V if op.V is not None else np.zeros((0, 0)), E=S ** 2 / data.nsnps,
                 epochs=op.epochs, trace=op.trace, em_trace=em_trace, perm=data.perm)
    if p.ld:  # src/Halko.cpp:338
        res.residuals = residuals(data, op.S, op.U, op.V.T)
    return res


def residuals(data: Data, S, U, VT) -> np.ndarray:
    """Data::write_residuals (src/Data.cpp:242-291) as the M x N float32 matrix
    of the .residuals file (rows in file order).  In-core: the matrix as the
    run left it (imputed and standardized as the run made it), minus U S V^T
    with --ld-stats 0, each SNP centered over samples, cast to float; with the
    in-core shuffle SNP column ib goes to row perm[ib].  Out-of-core: every
    block is read again centered and unscaled, without imputation
    (read_block_initial(.., false)), and the same follows."""
    p = data.params
    G = (data.G0 if p.out_of_core else data.G).copy()
    if p.ld_stats == 0:
        G -= (U * S[None, :]) @ VT
    G -= G.mean(axis=0, keepdims=True)
    out = np.empty((G.shape[1], G.shape[0]), dtype=np.float32)
    rows = data.perm if (p.perm and data.perm is not None) else np.arange(G.shape[1])
    out[rows] = G.T.astype(np.float32)
    return out


def write_residuals_file(path: str, R: np.ndarray) -> None:
    """The .residuals layout (src/Data.cpp:252-256): uint32 nsnps, uint32
    nsamples, then nsamples float32 values per SNP."""
    with open(path, "wb") as f:
        f.write(np.array([R.shape[0], R.shape[1]], dtype=np.uint32).tobytes())
        f.write(np.ascontiguousarray(R, dtype=np.float32).tobytes())


def read_residuals_file(path: str) -> np.ndarray:
    raw = np.fromfile(path, dtype=np.uint8)
    m, n = np.frombuffer(raw[:8].tobytes(), dtype=np.uint32)
    return np.frombuffer(raw[8:].tobytes(), dtype=np.float32).reshape(int(m), int(n))


def run_pca_full(data: Data) -> Result:
    """--svd 3 (src/Main.cpp:180-217): standardize_E, then the eigen-decomposition
    of G G^T / M (N <= M) or G^T G / M, the ncomp = min(k, N, M) largest
    eigenvalues (clamped at 0) as evals, svals = sqrt(evals M), the other side
    G^T U / s (or G V / s), flip_UV with U as the base."""
    p = data.params
    data.standardize_E()
    G = data.G
    N, M = G.shape
    ncomp = min(p.k, N, M)
    if N <= M:
        w, Q = np.linalg.eigh(G @ G.T / M)
        idx = np.arange(len(w) - 1, len(w) - 1 - ncomp, -1)
        evals = np.maximum(w[idx], 0.0)
        U = Q[:, idx]
        S = np.sqrt(evals * M)
        V = G.T @ U
        V[:, S > 0] /= S[S > 0]
    else:
        w, Q = np.linalg.eigh(G.T @ G / M)
        idx = np.arange(len(w) - 1, len(w) - 1 - ncomp, -1)
        evals = np.maximum(w[idx], 0.0)
        V = Q[:, idx]
        S = np.sqrt(evals * M)
        U = G @ V
        U[:, S > 0] /= S[S > 0]
    U, V = flip_uv(U, V, True)
    return Result(U=U, S=S, V=V, E=evals, epochs=1)


def run_pca_arnoldi(data: Data) -> Result:
    """--svd 0, in-core without EMU (src/Arnoldi.cpp:48-118): Spectra's
    PartialSVDSolver on the standardized G (unscaled with --ld) converged to
    --itol: the k largest eigenpairs of G G^T (N <= M; G^T G otherwise), so
    U, S = sqrt(eigenvalues), V = G^T U / S (or the tall mirror), E = S^2 / M;
    no flip_UV.  Restated as the exact eigenpairs (what the converged IRAM
    fixes); column signs are the solver's and are compared sign-aligned."""
    p = data.params
    if not p.ld:
        data.standardize_E()
    G = data.G
    N, M = G.shape
    k = p.k
    if N <= M:
        w, Q = np.linalg.eigh(G @ G.T)
        idx = np.arange(len(w) - 1, len(w) - 1 - k, -1)
        S = np.sqrt(np.maximum(w[idx], 0.0))
        U = Q[:, idx]
        V = G.T @ U / S
    else:
        w, Q = np.linalg.eigh(G.T @ G)
        idx = np.arange(len(w) - 1, len(w) - 1 - k, -1)
        S = np.sqrt(np.maximum(w[idx], 0.0))
        V = Q[:, idx]
        U = G @ V / S
    res = Result(U=U, S=S, V=V, E=S ** 2 / M, epochs=1)
    if p.ld:  # src/Arnoldi.cpp:117
        res.residuals = residuals(data, S, U, V.T)
    return res


def run_pca(data: Data) -> Result:
    """The dispatch of src/Main.cpp:176-217 by --svd."""
    if data.params.svd == 3:
        return run_pca_full(data)
    if data.params.svd == 0:
        return run_pca_arnoldi(data)
    return run_pca_with_halko(data)


def run_plink(prefix: str, params: Params) -> Result:
    packed, n, _ = read_plink(prefix)
    return run_pca(Data.from_packed(packed, n, params))


# ---------------------------------------------------------------- outputs
def fmt6(x: float) -> str:
    """Eigen IOFormat(6, ...) element: ostream with precision 6 (src/Data.cpp:215)."""
    return f"{x:.6g}"


def write_eigs_files(res: Result, data: Data, out: str, printv: bool = False) -> None:
    """src/Data.cpp:211-240 (.sigvals/.eigvals/.eigvecs[/.loadings])."""
    with open(out + ".sigvals", "w") as f:
        f.write(f"#{res.U.shape[0]},{res.V.shape[0]}\n")
        f.write("\n".join(fmt6(v) for v in res.S) + "\n")
    with open(out + ".eigvals", "w") as f:
        f.write("\n".join(fmt6(v) for v in res.E) + "\n")
    with open(out + ".eigvecs", "w") as f:
        f.write("\n".join("\t".join(fmt6(v) for v in row) for row in res.U) + "\n")
    if printv:
        V = res.V
        if data.params.perm and data.perm is not None:
            V = np.empty_like(res.V)
            V[data.perm] = res.V  # original SNP perm[i] <- permuted row i
        with open(out + ".loadings", "w") as f:
            f.write("\n".join("\t".join(fmt6(v) for v in row) for row in V) + "\n")


# ---------------------------------------------------------------- exact oracle
def dense_pca(data: Data, k: int):
    """Exact top-k PCs of the standardized matrix (the reference's --svd 3
    quantity, src/Main.cpp:180-217): U (N x k), S, E = S^2/M, V (M x k)."""
    U, s, Vt = np.linalg.svd(data.G, full_matrices=False)
    return U[:, :k], s[:k], s[:k] ** 2 / data.nsnps, Vt[:k].T


# ---------------------------------------------------------------- comparison metrics
def sign_align(ref: np.ndarray, x: np.ndarray) -> np.ndarray:
    sgn = np.sign(np.sum(ref * x, axis=0))
    sgn[sgn == 0] = 1
    return x * sgn[None, :]


def pc_correlations(ref: np.ndarray, x: np.ndarray) -> np.ndarray:
    """|Pearson correlation| of each PC column pair."""
    out = []
    for i in range(ref.shape[1]):
        a = ref[:, i] - ref[:, i].mean()
        b = x[:, i] - x[:, i].mean()
        out.append(abs(float(a @ b) / (np.linalg.norm(a) * np.linalg.norm(b) + 1e-300)))
    return np.array(out)


def principal_angles(ref: np.ndarray, x: np.ndarray) -> np.ndarray:
    qa = np.linalg.qr(ref)[0]
    qb = np.linalg.qr(x)[0]
    s = np.linalg.svd(qa.T @ qb, compute_uv=False)
    return np.arccos(np.clip(s, -1.0, 1.0))


def read_matrix(path: str) -> np.ndarray:
    rows = []
    with open(path) as f:
        for line in f:
            line = line.strip()
            if not line or line.startswith("#"):
                continue
            rows.append([float(t) for t in line.split("\t")])
    return np.array(rows)
