"""Writes PLINK 2 .pgen/.pvar/.psam file sets from PLINK 1 .bed/.bim/.fam sets
(test infrastructure: the inputs of the PGEN reader's tests and golden runs).

The storage modes and record types follow the PGEN layout that pgenlib reads
(external/pgenlib/include/pgenlib_read.cc: PgfiInitPhase1 :684-957 for the
three header forms, PgfiInitPhase2Ex :1094-1480 for the variable-width index,
ReadGenovecSubsetUnsafe :2837-2900 and ParseNonLdGenovecSubsetUnsafe
:2701-2731 for the main-track record types, ParseDifflistHeader /
ParseAndSaveDifflist :2177-2267 for difference lists, ParseOnebitUnsafe
:2597-2681, ParseDosage16 :7185 for the unconditional dosage track).  The
writer is pinned by the reference binary itself: pgenlib, compiled from the
reference sources, reads every file written here, and its PCA of the file
must equal its PCA of the .bed it came from (tests/golden/make_golden.py).

Genotype codes: a .bed code (00 two A1 copies, 01 missing, 10 one copy, 11 no
copy, A1 the counted allele) becomes the PGEN hardcall ALT count 2, 3
(missing), 1, 0, with ALT = A1 in the .pvar (src/Utils.cpp:410 exposes ALT as
the .bim A1).

Variable-width files use every main-track record type: 0 (2-bit array),
1 (two common genotypes as a bit array plus a difference list), 2 and 3
(difference list against the last earlier variant that is not LD-compressed;
3 then swaps 0 and 2), 4 / 6 / 7 (one genotype everywhere plus a difference
list), 5 (all homozygous reference, empty record).  `policy` picks among the
types that fit (a difference list holds at most N / 8 entries).
"""
from __future__ import annotations

import struct
from pathlib import Path

import numpy as np

VBLOCK = 65536
BED_TO_ALT = np.array([2, 3, 1, 0], dtype=np.uint8)  # .bed code -> ALT count (3 = missing)


def bed_to_geno(rows: np.ndarray, n: int) -> np.ndarray:
    """Packed .bed rows (M x ceil(n/4)) -> M x n ALT counts (3 = missing)."""
    m = rows.shape[0]
    codes = np.empty((m, n), dtype=np.uint8)
    for e in range(4):
        cols = np.arange(e, n, 4)
        codes[:, cols] = (rows[:, cols // 4] >> (2 * e)) & 3
    return BED_TO_ALT[codes]


def pack2(vals: np.ndarray) -> bytes:
    """2-bit entries, entry e at bits 2(e % 4) of byte e / 4."""
    v = np.zeros(((len(vals) + 3) // 4) * 4, dtype=np.uint8)
    v[: len(vals)] = vals
    v = v.reshape(-1, 4)
    return (v[:, 0] | (v[:, 1] << 2) | (v[:, 2] << 4) | (v[:, 3] << 6)).astype(np.uint8).tobytes()


def varint(x: int) -> bytes:
    out = bytearray()
    while True:
        b = x & 127
        x >>= 7
        if x:
            out.append(b | 128)
        else:
            out.append(b)
            return bytes(out)


def sample_id_bytes(n: int) -> int:
    return 1 + (n.bit_length() - 1) // 8


def difflist(ids: np.ndarray, genos: np.ndarray, n: int) -> bytes:
    """A difference list: varint length; first sample ID of every group of 64
    (fixed width); per group but the last, the byte count of its deltas minus
    63; the 2-bit genotypes; the deltas inside each group as varints."""
    ln = len(ids)
    out = bytearray(varint(ln))
    if ln == 0:
        return bytes(out)
    sb = sample_id_bytes(n)
    groups = [ids[g:g + 64] for g in range(0, ln, 64)]
    deltas = [b"".join(varint(int(d)) for d in np.diff(g)) for g in groups]
    for g in groups:
        out += int(g[0]).to_bytes(sb, "little")
    for d in deltas[:-1]:
        out.append(len(d) - 63)
    out += pack2(genos)
    for d in deltas:
        out += d
    return bytes(out)


def invert(g: np.ndarray) -> np.ndarray:
    """Swaps ALT counts 0 and 2 (missing and 1 stay)."""
    out = g.copy()
    out[g == 0] = 2
    out[g == 2] = 0
    return out


def encode_variant(g: np.ndarray, ldbase: np.ndarray | None, j: int, policy: str):
    """(vrtype, main-track bytes) of one variant's ALT counts g."""
    n = len(g)
    lim = n // 8
    cand = {0: None}
    counts = np.bincount(g, minlength=4)
    if counts[0] == n:
        cand[5] = None
    for base, vt in ((0, 4), (2, 6), (3, 7)):
        if n - counts[base] <= lim:
            cand[vt] = base
    order = np.argsort(-counts, kind="stable")
    lo, hi = sorted(int(x) for x in order[:2])
    if counts[lo] + counts[hi] >= n - lim and counts[hi] > 0:
        cand[1] = (lo, hi)
    if ldbase is not None:
        if int(np.count_nonzero(g != ldbase)) <= lim:
            cand[2] = None
        if int(np.count_nonzero(invert(g) != ldbase)) <= lim:
            cand[3] = None
    if policy == "raw":
        vt = 0
    elif policy == "plink2":
        # smallest record wins, as a compressing writer would choose
        vt = min(cand, key=lambda t: len(encode_as(t, cand[t], g, ldbase)))
    else:  # "cycle": rotate through the types that fit, so every one appears
        keys = sorted(cand)
        vt = keys[j % len(keys)]
    return vt, encode_as(vt, cand[vt], g, ldbase)


def encode_as(vt, arg, g, ldbase) -> bytes:
    n = len(g)
    if vt == 0:
        return pack2(g)
    if vt == 5:
        return b""
    if vt in (4, 6, 7):
        ids = np.nonzero(g != arg)[0]
        return difflist(ids, g[ids], n)
    if vt == 1:
        lo, hi = arg
        bits = (g == hi).astype(np.uint8)
        ids = np.nonzero((g != lo) & (g != hi))[0]
        return bytes([lo * 4 + (hi - lo)]) + np.packbits(bits, bitorder="little").tobytes() + \
            difflist(ids, g[ids], n)
    target = g if vt == 2 else invert(g)
    ids = np.nonzero(target != ldbase)[0]
    return difflist(ids, target[ids], n)


def write_pgen(path: str, geno: np.ndarray, mode: str = "var", policy: str = "cycle", len_bytes: int = 0,
               nonref: int = 0, dosage: bool = False) -> None:
    """geno: M x N ALT counts (3 = missing).  mode: 'plink1' (a .bed payload
    behind the 6c 1b 01 header), 'fixed' (mode 0x02, 2-bit arrays), 'var'
    (mode 0x10, variable-width records).  len_bytes: bytes per record length
    (0 = the fewest that fit).  nonref: header bits 6-7 (3 stores one flag per
    variant).  dosage: 8-bit vrtypes and an unconditional dosage track (the
    hardcall x 16384, 65535 missing) after every main track."""
    m, n = geno.shape
    f = bytearray(b"l\x1b")
    if mode == "plink1":
        inv = np.array([3, 2, 0, 1], dtype=np.uint8)  # ALT count -> .bed code
        f.append(1)
        for j in range(m):
            f += pack2(inv[geno[j]])
        Path(path).write_bytes(bytes(f))
        return
    if mode == "fixed":
        f.append(2)
        f += struct.pack("<II", m, n)
        f.append(nonref << 6)
        if nonref == 3:
            f += np.packbits(np.arange(m) % 3 == 0, bitorder="little").tobytes()
        for j in range(m):
            f += pack2(geno[j])
        Path(path).write_bytes(bytes(f))
        return
    recs, vrtypes = [], []
    ldbase = None
    for j in range(m):
        g = geno[j]
        vt, body = encode_variant(g, ldbase if j % VBLOCK else None, j, policy)
        if vt not in (2, 3):
            ldbase = g
        if dosage:
            d = g.astype(np.uint32) * 16384
            d[g == 3] = 65535
            body += d.astype("<u2").tobytes()
            vt |= 0x40
        recs.append(body)
        vrtypes.append(vt)
    maxlen = max(len(r) for r in recs)
    lb = len_bytes or next(b for b in (1, 2, 3, 4) if maxlen < (1 << (8 * b)))
    assert maxlen < (1 << (8 * lb))
    ctrl = (lb - 1) | (4 if dosage else 0) | (nonref << 6)
    f.append(0x10)
    f += struct.pack("<II", m, n)
    f.append(ctrl)
    nvb = (m + VBLOCK - 1) // VBLOCK
    fpos_at = len(f)
    f += b"\0" * (8 * nvb)
    for b in range(nvb):
        j0, j1 = b * VBLOCK, min(m, (b + 1) * VBLOCK)
        vts = np.array(vrtypes[j0:j1], dtype=np.uint8)
        if dosage:
            f += vts.tobytes()
        else:
            v = np.zeros(((j1 - j0 + 1) // 2) * 2, dtype=np.uint8)
            v[: j1 - j0] = vts
            f += (v[0::2] | (v[1::2] << 4)).astype(np.uint8).tobytes()
        for r in recs[j0:j1]:
            f += len(r).to_bytes(lb, "little")
        if nonref == 3:
            f += np.packbits(np.arange(j0, j1) % 3 == 0, bitorder="little").tobytes()
    pos = len(f)
    for b in range(nvb):
        struct.pack_into("<Q", f, fpos_at + 8 * b, pos)
        pos += sum(len(r) for r in recs[b * VBLOCK:(b + 1) * VBLOCK])
    for r in recs:
        f += r
    Path(path).write_bytes(bytes(f))


def write_pvar_psam(bfile: str, prefix: str, fid: bool = True) -> None:
    """.pvar from the .bim (REF = A2, ALT = A1; '##' and '#CHROM' header
    lines) and .psam from the .fam ('#FID IID SEX', or '#IID SEX' with the
    IIDs alone when fid is False)."""
    pv = ["##fileformat=PVARv1.0", "##source=pcaone_amd test writer", "#CHROM\tPOS\tID\tREF\tALT"]
    for line in Path(bfile + ".bim").read_text().splitlines():
        t = line.split()
        pv.append(f"{t[0]}\t{t[3]}\t{t[1]}\t{t[5]}\t{t[4]}")
    Path(prefix + ".pvar").write_text("\n".join(pv) + "\n")
    ps = ["#FID\tIID\tSEX" if fid else "#IID\tSEX"]
    for line in Path(bfile + ".fam").read_text().splitlines():
        t = line.split()
        ps.append(f"{t[0]}\t{t[1]}\tNA" if fid else f"{t[1]}\tNA")
    Path(prefix + ".psam").write_text("\n".join(ps) + "\n")


def bed_to_pgen(bfile: str, prefix: str, fid: bool = True, **kw) -> None:
    """A .bed/.bim/.fam set as .pgen/.pvar/.psam (write_pgen's options)."""
    raw = Path(bfile + ".bed").read_bytes()
    n = sum(1 for _ in Path(bfile + ".fam").read_text().splitlines())
    rows = np.frombuffer(raw[3:], dtype=np.uint8).reshape(-1, (n + 3) // 4)
    write_pgen(prefix + ".pgen", bed_to_geno(rows, n), **kw)
    write_pvar_psam(bfile, prefix, fid=fid)
