"""Regenerates the golden fixtures under tests/golden/ from the reference binary.

Run in the build container (the reference sources are needed to build
oracle/_ref/pcaone_ref):

    python oracle/build_oracle.py && python tests/golden/make_golden.py

Inputs are synthetic Balding-Nichols PLINK sets.  The small ones are committed
under tests/golden/data/; the 2,000 x 20,000 one is regenerated from its spec
and checked against the recorded SHA-256.  For every case the reference's
own output files (.eigvals/.sigvals/.eigvecs[/.loadings/.mbim]) are copied
and meta.json records the command, the epoch count and the per-epoch
convergence trace parsed from the log.
"""
from __future__ import annotations

import hashlib
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(HERE))

import numpy as np  # noqa: E402

from pcaone_amd import synth  # noqa: E402

REF = ROOT / "oracle" / "_ref" / "pcaone_ref"
# the complete reference program (oracle/ref_full.mk): the --svd 0 and --svd 3 branches
REF_FULL = ROOT / "oracle" / "_ref" / "pcaone_full"
DATA = HERE / "data"

INPUTS = {
    "f1": dict(spec=synth.Spec(seed=2001, nsamples=200, nsnps=2000, npops=3), commit=True),
    "f4": dict(spec=synth.Spec(seed=2004, nsamples=300, nsnps=3000, npops=3, miss_rate=0.1), commit=True),
    "cfg1": dict(spec=synth.config(1), commit=False),
    # edge cases: ragged N (150), monomorphic, all-missing and single-genotype SNPs
    "f5": dict(spec=synth.Spec(seed=2005, nsamples=150, nsnps=1200, npops=3, miss_rate=0.05), commit=True,
               edit="edge"),
    # rank deficient: 12 distinct samples, each present 16 times (N = 192), so the
    # standardized matrix has rank 11 (centering removes one)
    "f6": dict(spec=synth.Spec(seed=2006, nsamples=12, nsnps=1500, npops=3), commit=True, edit="dup16"),
    # small input for the residual-matrix (-D/--ld) files: 64 x 500, 5 % missing
    "f7": dict(spec=synth.Spec(seed=2007, nsamples=64, nsnps=500, npops=3, miss_rate=0.05), commit=True),
    # more samples than SNPs (256 x 150): the tall branches of --svd 0 and --svd 3
    "f8": dict(spec=synth.Spec(seed=2008, nsamples=256, nsnps=150, npops=4), commit=True),
    # 1,008 SNPs in 64 windows of 16: window 63 starts at SNP 1,008 = M, an empty window
    "f9": dict(spec=synth.Spec(seed=7, nsamples=150, nsnps=1008, npops=3), commit=True),
    # LD structure for the .pgen record types: SNP 2i + 1 repeats SNP 2i with
    # 0 / 5 / 40 / 80 entries changed, every third pair with ALT and REF
    # swapped; N = 700 (two-byte sample indices, difference lists of two groups)
    "f10": dict(spec=synth.Spec(seed=2010, nsamples=700, nsnps=1200, npops=3, miss_rate=0.03), commit=True,
                edit="ld"),
}


def dup_samples(rows, n, reps):
    """Packed rows of n samples -> rows of n * reps samples, sample i a copy of
    sample i % n."""
    m = rows.shape[0]
    codes = np.zeros((m, n), dtype=np.uint8)
    for i in range(n):
        codes[:, i] = (rows[:, i // 4] >> (2 * (i % 4))) & 3
    full = np.tile(codes, (1, reps))
    N = n * reps
    out = np.zeros((m, (N + 3) // 4), dtype=np.uint8)
    for i in range(N):
        out[:, i // 4] |= full[:, i] << (2 * (i % 4))
    return out


def edge_rows(rows, n):
    """Overwrites a few SNPs of a packed matrix (rows: M x ceil(n/4) bytes) with
    the reference's degenerate cases: all code 00 (F = 1), all code 11 (F = 0),
    all missing (no count, F = 0, src/FilePlink.cpp:51-52), one genotype among
    missing entries (F = 0.5, and F = 1), all heterozygous (centered to zero)."""
    def fill(j, codes):
        row = [0] * rows.shape[1]
        for i, c in enumerate(codes):
            row[i // 4] |= (c & 3) << (2 * (i % 4))
        rows[j] = row
    fill(3, [0] * n)
    fill(10, [3] * n)
    fill(17, [1] * n)
    fill(40, [1] * 70 + [2] + [1] * (n - 71))
    fill(41, [1] * 5 + [0] + [1] * (n - 6))
    fill(100, [2] * n)
    return rows

def ld_rows(rows, n, seed=10):
    """Packed rows -> rows whose odd SNPs copy the SNP before them with a few
    entries redrawn (and ALT/REF swapped for every third pair), plus SNPs that
    are homozygous everywhere or missing nearly everywhere."""
    import pgen_writer as W
    rng = np.random.default_rng(seed)
    g = W.bed_to_geno(rows, n)
    for i in range(g.shape[0] // 2):
        a = g[2 * i].copy()
        c = (0, 5, 40, 80)[i % 4]
        idx = rng.choice(n, size=c, replace=False)
        a[idx] = (a[idx] + 1 + rng.integers(0, 3, size=c)) % 4
        g[2 * i + 1] = W.invert(a) if i % 3 == 1 else a
    g[20] = 0
    g[21] = 2
    g[22] = 3
    g[22, :30] = 1
    g[23] = 2
    g[23, rng.choice(n, size=70, replace=False)] = 0
    g[26] = 3
    g[26, rng.choice(n, size=12, replace=False)] = 2
    inv = np.array([3, 2, 0, 1], dtype=np.uint8)  # ALT count -> .bed code
    return np.stack([np.frombuffer(W.pack2(inv[r]), dtype=np.uint8) for r in g])


CASES = {
    "f1_d1": ("f1", "-d 1 -k 5 --maxp 7 -V"),
    "f1_d1_p0": ("f1", "-d 1 -k 5 --maxp 0"),
    "f1_d1_unif": ("f1", "-d 1 -k 5 --maxp 3 --rand 0 --seed 7 --tol-rsvd 0"),
    "f1_d1_maf": ("f1", "-d 1 -k 4 --maf 0.1 --maxp 5 -V"),
    "f1_d2": ("f1", "-d 2 -w 16 -k 5 --maxp 7 -V"),
    "f1_d2_noshuf": ("f1", "-d 2 -w 16 -k 5 --maxp 7 -V -S"),
    "f4_d1_miss": ("f4", "-d 1 -k 3 --maxp 5 -V"),
    "f4_emu_d1": ("f4", "-d 1 -k 3 --emu --maxiter 5 -V"),
    "f4_emu_d2": ("f4", "-d 2 -w 16 -k 3 --emu --maxiter 5 -V"),
    "cfg1_d1": ("cfg1", "-d 1 -k 10 --oversamples 10 --maxp 7 --tol-rsvd 0"),
    "cfg1_d2": ("cfg1", "-d 2 -w 64 -k 10 --maxp 7 --tol-rsvd 0"),
    # out-of-core (-m GB): 3 blocks; 10 blocks with the minimum-RAM branch of the plan
    "f1_d1_ooc": ("f1", "-d 1 -k 5 --maxp 7 -V -m 0.002"),
    "f1_d1_ooc_min": ("f1", "-d 1 -k 5 --maxp 7 --tol-rsvd 0 -m 0.0003"),
    # winSVD over 48 blocks (band factor 3) in permute_plink order; 32 blocks unshuffled
    "f1_d2_ooc": ("f1", "-d 2 -w 16 -k 5 --maxp 7 -V -m 0.00068"),
    "f1_d2_ooc_noshuf": ("f1", "-d 2 -w 16 -k 5 --maxp 7 -S -m 0.0007"),
    # EMU with per-pass block imputation; winSVD over 64 blocks (band factor 4),
    # 3000 SNPs in 16 bands (the shuffle's remainder path)
    "f4_emu_d1_ooc": ("f4", "-d 1 -k 3 --emu --maxiter 5 -V -m 0.002"),
    "f4_emu_d2_ooc": ("f4", "-d 2 -w 16 -k 3 --emu --maxiter 5 -V -m 0.0009"),
    # degenerate SNPs (monomorphic, all missing, one genotype) through each driver
    "f5_d1": ("f5", "-d 1 -k 4 --maxp 6 -V"),
    "f5_d2": ("f5", "-d 2 -w 16 -k 4 --maxp 7 -V"),
    "f5_emu": ("f5", "-d 1 -k 3 --emu --maxiter 4 -V"),
    "f5_emu_d2": ("f5", "-d 2 -w 16 -k 3 --emu --maxiter 4 -V"),
    "f5_maf": ("f5", "-d 1 -k 4 --maf 0.02 --maxp 5 -V"),
    # rank 11 < L = 16: the Gram matrices of G and H are singular
    "f6_d1": ("f6", "-d 1 -k 8 --maxp 7 -V"),
    "f6_d2": ("f6", "-d 2 -w 16 -k 8 --maxp 7 -V"),
    # L = 10 just under the rank
    "f6_near": ("f6", "-d 1 -k 5 --oversamples 5 --maxp 7 --tol-rsvd 0 -V"),
    # -D/--ld: the residual matrix (src/Data.cpp:242-291) after an unscaled run
    "f7_ld": ("f7", "-d 1 -k 3 --maxp 7 -D"),
    "f7_ld_std": ("f7", "-d 1 -k 3 --maxp 7 -D --ld-stats 1"),
    "f7_ld_d2": ("f7", "-d 2 -w 8 -k 3 --maxp 7 -D"),
    "f7_ld_emu": ("f7", "-d 1 -k 3 --emu --maxiter 4 -D"),
    "f7_ld_ooc": ("f7", "-d 2 -w 8 -k 3 --maxp 7 -D -m 0.0002"),
    # --svd 3: exact PCA (src/Main.cpp:180-217); --svd 0: IRAM (src/Arnoldi.cpp:48-118)
    "f1_d3": ("f1", "-d 3 -k 5 -V"),
    "f4_d3": ("f4", "-d 3 -k 3 -V"),
    "f8_d3": ("f8", "-d 3 -k 4 -V"),
    "f1_d0": ("f1", "-d 0 -k 4 -V"),
    "f8_d0": ("f8", "-d 0 -k 3 -V"),
    "f7_d0_ld": ("f7", "-d 0 -k 3 -D"),
    # winSVD whose last window is empty (src/Halko.cpp:191-194 with a zero width)
    "f9_d2_empty": ("f9", "-d 2 -w 64 -k 4 --maxp 7 --tol-rsvd 0 -V"),
    # PLINK 2 input (-p/--pgen, src/FilePgen.cpp): the .pgen written from the
    # input by pgen_writer with the options given (not committed; its SHA-256
    # is recorded, and the tests write it again from the committed .bed)
    "f10_pg_d1": ("f10", "-d 1 -k 5 --maxp 7 -V", dict(mode="var", policy="cycle")),
    "f10_pg_d2": ("f10", "-d 2 -w 16 -k 5 --maxp 7 -V",
                  dict(mode="var", policy="plink2", nonref=3, len_bytes=2, fid=False)),
    "f1_pg_ooc": ("f1", "-d 2 -w 16 -k 5 --maxp 7 -V -m 0.00068 -n 4", dict(mode="var", policy="cycle")),
    "f1_pg_p1": ("f1", "-d 1 -k 5 --maxp 7 -V", dict(mode="plink1")),
    "f1_pg_dos": ("f1", "-d 1 -k 5 --maxp 7 --hardcall -V", dict(mode="var", policy="cycle", dosage=True,
                                                                   len_bytes=3)),
    "f4_pg_emu": ("f4", "-d 1 -k 3 --emu --maxiter 5 -V", dict(mode="fixed", fid=False)),
    "f4_pg_emu_ooc": ("f4", "-d 2 -w 16 -k 3 --emu --maxiter 5 -V -m 0.0009 -n 4", dict(mode="var", policy="plink2")),
    "f5_pg_maf": ("f5", "-d 1 -k 4 --maf 0.02 --maxp 5 -V", dict(mode="var", policy="plink2")),
}

EPOCH_RE = re.compile(r"running of epoch =\s*(\d+), diff = ([-0-9.e+]+)")
STOP_RE = re.compile(r"stops at epoch =\s*(\d+)")
EM_RE = re.compile(r"estimated iter =\s*(\d+), diff = ([-0-9.e+]+)")
# Data::prepare's out-of-core plan as the reference logs it (src/Data.cpp:60-61, 73-74)
PLAN_RE = re.compile(r"blocksize =\s*(\d+), nblocks =\s*(\d+), factor =\s*(\d+)")


def sha256(path: Path) -> str:
    return hashlib.sha256(path.read_bytes()).hexdigest()


def make_inputs(tmp: Path) -> dict:
    DATA.mkdir(exist_ok=True)
    out = {}
    for name, d in INPUTS.items():
        spec = d["spec"]
        prefix = (DATA if d["commit"] else tmp) / name
        synth.write_plink(spec, str(prefix))
        if d.get("edit") == "edge":
            bed = Path(str(prefix) + ".bed")
            raw = bed.read_bytes()
            rows = np.frombuffer(raw[3:], dtype=np.uint8).reshape(spec.nsnps, -1).copy()
            bed.write_bytes(raw[:3] + edge_rows(rows, spec.nsamples).tobytes())
        if d.get("edit") == "ld":
            bed = Path(str(prefix) + ".bed")
            raw = bed.read_bytes()
            rows = np.frombuffer(raw[3:], dtype=np.uint8).reshape(spec.nsnps, -1)
            bed.write_bytes(raw[:3] + ld_rows(rows, spec.nsamples).tobytes())
        if d.get("edit", "").startswith("dup"):
            reps = int(d["edit"][3:])
            bed = Path(str(prefix) + ".bed")
            raw = bed.read_bytes()
            rows = np.frombuffer(raw[3:], dtype=np.uint8).reshape(spec.nsnps, -1)
            bed.write_bytes(raw[:3] + dup_samples(rows, spec.nsamples, reps).tobytes())
            Path(str(prefix) + ".fam").write_text(
                "".join(f"F{i} I{i} 0 0 0 -9\n" for i in range(spec.nsamples * reps)))
        out[name] = dict(prefix=str(prefix), sha256=sha256(Path(str(prefix) + ".bed")), spec=spec.__dict__,
                         committed=d["commit"])
    return out


def run_case(name: str, inp: dict, args: str, tmp: Path, pgen: dict | None = None) -> None:
    dest = HERE / name
    if dest.exists():
        shutil.rmtree(dest)
    dest.mkdir()
    outp = tmp / name
    full = "-d 0" in args or "-d 3" in args or pgen is not None
    if pgen is not None:
        import pgen_writer as W
        pin = str(tmp / (name + "_in"))
        W.bed_to_pgen(inp["prefix"], pin, **pgen)
        cmd = [str(REF_FULL), "-p", pin, *args.split(), "-o", str(outp), "-v", "1"]
    else:
        cmd = [str(REF_FULL if full else REF), "-b", inp["prefix"], *args.split(), "-o", str(outp), "-v", "1"]
    env = dict(os.environ, OMP_NUM_THREADS=os.environ.get("OMP_NUM_THREADS", "8"))
    if "-n " in args:
        # FilePgen keeps one reader state per -n thread and indexes them by the
        # OpenMP thread number (src/FilePgen.hpp:30-36, src/FilePgen.cpp:165-168),
        # so the OpenMP team must not be larger than -n
        env["OMP_NUM_THREADS"] = args.split("-n ")[1].split()[0]
    res = subprocess.run(cmd, capture_output=True, text=True, env=env, check=True)
    log = res.stdout
    trace = [float(m.group(2)) for m in EPOCH_RE.finditer(log)]
    stop = STOP_RE.findall(log)
    em = [float(m.group(2)) for m in EM_RE.finditer(log)]
    plan = [[int(g) for g in m.groups()] for m in PLAN_RE.finditer(log)]
    sufs = [".eigvals", ".sigvals", ".eigvecs", ".loadings", ".mbim", ".residuals"]
    if inp["committed"]:
        sufs.append(".eigvecs2")  # the small inputs also keep the plink2-style file (src/Utils.cpp:225-238)
    for suf in sufs:
        src = Path(str(outp) + suf)
        if src.exists():
            shutil.copy(src, dest / suf[1:])
    meta = dict(input=inp_name_of(inp), args=args, trace=trace, stop_epochs=[int(s) for s in stop], em_trace=em)
    if pgen is not None:
        meta["pgen"] = dict(pgen, sha256=sha256(Path(pin + ".pgen")))
    if plan:
        meta["block_plan"] = plan[-1]  # [blocksize, nblocks, factor] after any winSVD adjustment
    (dest / "meta.json").write_text(json.dumps(meta, indent=1) + "\n")
    print(f"{name}: {len(trace)} checked epochs, stops {stop}, em {len(em)}")


_names = {}


def inp_name_of(inp):
    return _names[inp["prefix"]]


def main() -> None:
    if not REF.exists():
        raise SystemExit(f"{REF} missing: python oracle/build_oracle.py")
    with tempfile.TemporaryDirectory() as td:
        tmp = Path(td)
        inputs = make_inputs(tmp)
        for n, inp in inputs.items():
            _names[inp["prefix"]] = n
        manifest = {n: dict(sha256=i["sha256"], spec=i["spec"], committed=i["committed"]) for n, i in inputs.items()}
        (HERE / "inputs.json").write_text(json.dumps(manifest, indent=1) + "\n")
        only = set(sys.argv[1:])  # optional: regenerate only the named cases
        for name, (inp, args, *pg) in CASES.items():
            if not only or name in only:
                run_case(name, inputs[inp], args, tmp, pg[0] if pg else None)


if __name__ == "__main__":
    main()
