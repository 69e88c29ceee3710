#!/usr/bin/env python3
"""Benchmark: time to the top-k PCs of PCAone's single-pass randomized SVD
(--svd 1) on one MI355X, FP8 form (the headline; BF16 hi/lo and INT8 beside
it) against the plain FP32 form at equal error.

Workload (the precision form): 50,000 samples x 1,000,000 SNPs of synthetic
Balding-Nichols genotypes (8 populations, Fst 0.1, p ~ U(0.05, 0.5)),
generated straight into HBM as packed 2-bit PLINK codes; k = 40 with
oversampling 20 raised to 40 by the reference rule (L = 80); 7 power
iterations forced with --tol-rsvd 0 (8 passes over the data).  One step is
one complete randomized SVD (top 40 PCs) with the genotypes already resident.

Printed: one JSON line (rank 0).  Errors of the reduced forms are measured
against the FP64 form of the same run (same Omega) and against the same
workload in FP64 iterated to convergence.  The CPU
baseline is the reference's own binary (oracle/_ref/pcaone_ref, built from
the reference sources) on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "time-to-k PCs (s) + HBM/MFMA % of roofline, FP8/BF16 vs FP32 at equal error"
PEAK_TFLOPS = {"bf16": 2500.0, "bf16x1": 2500.0, "bf16mix": 2500.0, "fp16": 2500.0, "fp16mix": 2500.0,
               "int8": 5000.0, "fp8": 5000.0, "fp32": 157.3, "fp64": 78.6}  # dense MFMA (int8: TOP/s)
PEAK_HBM_GBS = 8000.0
FORM_TEXT = {
    "bf16": "BF16 MFMA, exact genotypes, hi/lo BF16 split of Omega and G, FP32 accumulation",
    "fp16": "FP16 MFMA, exact genotypes, one FP16 piece of Omega and G, FP32 accumulation",
    "bf16mix": "BF16 MFMA, exact genotypes, one BF16 piece of Omega, hi/lo of G, FP32 accumulation",
    "fp16mix": "FP16 (G = A Omega) and BF16 hi/lo (H = A^T G) MFMA, exact genotypes, FP32 accumulation",
    "int8": "int8 MFMA, exact genotypes (v/2), Omega and G as per-column 16-bit fixed point split into two int8 "
            "pieces, exact int32 accumulation",
    "fp8": "FP8 E4M3 MFMA (f8f6f4), exact genotypes, Omega and G scaled per column by a power of two and split "
           "into three E4M3 pieces, FP32 accumulation",
}
DTYPE = {"bf16": "bf16", "bf16x1": "bf16", "bf16mix": "bf16", "fp16": "fp16", "fp16mix": "fp16", "int8": "int8",
         "fp8": "fp8"}
# operand pieces per value of W (Omega or G): executed MFMA work = pieces x algorithmic
PIECES = {"bf16": 2, "bf16x1": 1, "bf16mix": 2, "fp16": 1, "fp16mix": 2, "int8": 2, "fp8": 3, "fp32": 1, "fp64": 1}
SIDE_FORMS = ("bf16", "fp8", "int8")  # reported beside the headline (when not the headline) with their errors


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def errors(ref, x):
    from oracle import pcaone_oracle as O  # checker only
    rel = np.abs(x.E - ref.E) / ref.E
    corr = O.pc_correlations(ref.U, x.U)
    ang = O.principal_angles(ref.U, x.U)
    return {"eig_rel_err_max": float(rel.max()), "pc_corr_min": float(corr.min()),
            "principal_angle_max_rad": float(ang.max())}


def bar_met(e) -> bool:
    """The accuracy bar against the FP64 form of the same run: eigenvalue
    relative error <= 1e-4 and every PC correlated >= 0.9999."""
    return bool(e["eig_rel_err_max"] <= 1e-4 and e["pc_corr_min"] >= 0.9999)


def recorded_traffic(N, M, L, form, product):
    """HBM bytes per launch of the GEMM from the committed counter summary
    (FETCH_SIZE + WRITE_SIZE, calibrated on a plain copy), or None when the
    summary was not taken on this workload."""
    for f in sorted((ROOT / "profiles").glob("packed_gemm_*_traffic_*.md")):
        for line in f.read_text().splitlines():
            if not line.startswith("traffic_bytes_per_launch "):
                continue
            kv = dict(t.split("=", 1) for t in line.split()[1:])
            if (int(kv["nsamples"]), int(kv["nsnps"]), int(kv["L"]), kv["form"]) == (N, M, L, form):
                return float(kv[product])
    return None


def _ref_run(ref, prefix, k, os_, maxp, cores, memory, out):
    """One run of the reference binary; (compute seconds, wall, reading)."""
    cmd = [str(ref), "-b", prefix, "-d", "1", "-k", str(k), "--oversamples", str(os_), "--maxp", str(maxp),
           "--tol-rsvd", "0", "-n", str(cores), "-o", out, "-v", "1"]
    if memory > 0:
        cmd += ["-m", str(memory)]
    env = dict(os.environ, OMP_NUM_THREADS=str(cores))
    t0 = time.perf_counter()
    res = subprocess.run(cmd, capture_output=True, text=True, env=env, timeout=900)
    wall = time.perf_counter() - t0
    if res.returncode != 0:
        raise RuntimeError(f"reference binary failed: {res.stderr[-500:]}")
    m_wall = re.search(r"total elapsed wall time:\s*([0-9.]+)", res.stdout)
    m_read = re.search(r"total elapsed reading time:\s*([0-9.]+)", res.stdout)
    t_wall = float(m_wall.group(1)) if m_wall else wall
    t_read = float(m_read.group(1)) if m_read else 0.0
    return t_wall, t_read


def cpu_baseline(args, spec_full, k, os_, maxp):
    """The reference binary on two bounded samples of the workload (the same N,
    args.cpu_snps and 2 x args.cpu_snps SNPs) in its out-of-core mode (-m
    args.cpu_memory GB: the full matrix would need 400 GB of RAM in-core), its
    wall time fitted as a + b M and evaluated at the full M.  The intercept
    holds the per-run work that does not grow with M (the L x N solve and SVD of
    every epoch)."""
    from pcaone_amd import synth
    ref = ROOT / "oracle" / "_ref" / "pcaone_ref"
    host_cores = os.cpu_count() or 1
    cores = min(16, host_cores)
    sizes = (args.cpu_snps, 2 * args.cpu_snps)
    pts = []
    with tempfile.TemporaryDirectory() as td:
        for m_s in sizes:
            spec = synth.Spec(**{**spec_full.__dict__, "nsnps": m_s})
            prefix = os.path.join(td, f"sample{m_s}")
            t0 = time.perf_counter()
            synth.write_plink(spec, prefix)
            log(f"cpu sample written ({spec.nsamples} x {m_s}) in {time.perf_counter() - t0:.1f}s")
            if ref.exists():
                t_wall, t_read = _ref_run(ref, prefix, k, os_, maxp, cores, args.cpu_memory, os.path.join(td, "out"))
                kind = "reference"
                what = ("oracle/_ref/pcaone_ref (PCAone v0.7.2 path built from its sources, Eigen FP64, OpenMP), "
                        f"-m {args.cpu_memory} (out-of-core: blocks re-read and decoded every pass)")
            else:
                from oracle import pcaone_oracle as O
                packed = synth.packed(spec)
                t0 = time.perf_counter()
                data = O.Data.from_packed(packed, spec.nsamples,
                                          O.Params(svd=1, k=k, oversamples=os_, maxp=maxp, tol=0))
                t_read = time.perf_counter() - t0
                O.run_pca_with_halko(data)
                t_wall = time.perf_counter() - t0
                kind = "port"
                what = "numpy FP64 restatement (oracle/pcaone_oracle.py), in-core"
            pts.append((m_s, t_wall, t_read))
            log(f"cpu {kind} {spec.nsamples} x {m_s}: wall {t_wall:.2f}s (reading {t_read:.2f}s)")
    (m1, w1, r1), (m2, w2, r2) = pts
    slope = (w2 - w1) / (m2 - m1)
    inter = w1 - slope * m1
    full = inter + slope * spec_full.nsnps
    return {"value": round(full, 1), "unit": "s", "cores": cores, "host_cores": host_cores, "kind": kind,
            "sample": (f"{spec_full.nsamples} samples x {m1} and x {m2} SNPs, same generator and flags (-d 1 -k {k} "
                       f"--oversamples {os_} --maxp {maxp} --tol-rsvd 0); wall {w1:.2f}s and {w2:.2f}s (reading "
                       f"{r1:.2f}s, {r2:.2f}s) fitted as a + b M (a = {inter:.2f}s, b = {slope * 1e3:.3f} ms per "
                       f"SNP) and evaluated at M = {spec_full.nsnps}; OMP_NUM_THREADS = {cores} of {host_cores} "
                       f"host cores; {what}")}


CONFIG_RUNS = {
    # the headline workload through the out-of-core schedule (only with --memory)
    3: dict(params=dict(svd=1, k=40, oversamples=20, maxp=7, tol=0.0), forms=("fp8", "bf16", "fp32", "fp64"),
            what="--svd 1, k=40, L=80, 7 power iterations"),
    # winSVD, 64 windows, 10 epochs (--maxp 9 --tol-rsvd 0), k = 20 (L = 40)
    2: dict(params=dict(svd=2, k=20, oversamples=10, bands=64, maxp=9, tol=0.0),
            forms=("fp32", "bf16", "int8", "fp8", "fp64"),
            what="winSVD (--svd 2 -w 64 --maxp 9 --tol-rsvd 0), k=20, L=40"),
    # EMU, 20 EM iterations (--emu --maxiter 20 --tol-em 0), --svd 1, k = 10 (L = 20)
    5: dict(params=dict(svd=1, k=10, oversamples=10, emu=1, maxiter=20, tolem=0.0),
            forms=("bf16", "fp32", "int8", "fp8", "fp64"),
            what="EMU (--emu -d 1 --maxiter 20 --tol-em 0), k=10, L=20"),
}


def batch_line(args) -> int:
    """Batched form: 256 cohorts of uneven size (N log-uniform in [500, 20000],
    M log-uniform in [10,000, 200,000], 2-6 populations), --svd 1, k = 10,
    oversamples 10, 5 power iterations (6 passes), all cohorts in one
    persistent launch per step inside one HIP graph; against the same cohorts
    run one after another through the single-matrix path."""
    from oracle import pcaone_oracle as O  # checker only
    from pcaone_amd import api, synth
    specs = synth.cohorts(args.cohorts)
    ctx = api.Context(0)
    t0 = time.perf_counter()
    ctx.batch_generate(specs)
    log(f"batch: generated {len(specs)} cohorts in {time.perf_counter() - t0:.1f}s")
    kw = dict(svd=1, k=10, oversamples=10, maxp=5, tol=0.0)
    L, passes = 20, 6
    flops = sum(4.0 * sp.nsamples * sp.nsnps * L * passes for sp in specs)
    out = {"config": 4, "workload": f"{len(specs)} cohorts, N 500-20000, M 10k-200k (log-uniform), --svd 1, k=10, "
                                    f"L=20, {passes} passes, one persistent launch per step, one HIP graph",
           "cohorts": len(specs), "total_samples": sum(sp.nsamples for sp in specs),
           "total_snps": sum(sp.nsnps for sp in specs), "algorithmic_flop": flops}
    res, secs = {}, {}
    forms = [f for f in (args.forms.split(",") if args.forms else ["bf16", "fp8", "fp32"]) if f != "fp64"] + ["fp64"]
    for form in forms:
        p = api.default_params(precision=form, **kw)
        r = ctx.batch_run(p, use_graph=True)  # first run: tuning + graph capture
        log(f"batch {form}: first run {r['info']['seconds']:.3f}s, tuned G {r['info']['wm_g']}x{r['info']['waves_g']}/"
            f"{r['info']['kchunk_g']} H {r['info']['wm_h']}x{r['info']['waves_h']}/{r['info']['kchunk_h']}")
        ts = []
        for _ in range(max(args.steps, 1) if form != "fp64" else 1):
            t0 = time.perf_counter()
            r = ctx.batch_run(p, use_graph=True)
            ts.append(time.perf_counter() - t0)
        assert int(r["info"]["failed"]) == 0
        secs[form] = float(np.mean(ts))
        res[form] = (r, [ctx.batch_get_u(c) for c in range(len(specs))])
        out[f"{form}_info"] = r["info"]
        log(f"batch {form}: {secs[form]:.4f}s per run ({flops / secs[form] / 1e12:.1f} TFLOP/s algorithmic)")
    out["seconds"] = {f: round(v, 4) for f, v in secs.items()}
    out["algorithmic_tflops"] = {f: round(flops / v / 1e12, 1) for f, v in secs.items()}
    r64, u64 = res["fp64"]
    errs = {}
    for f in res:
        if f == "fp64":
            continue
        r, us = res[f]
        rel = max(float(np.max(np.abs(r["E"][c] - r64["E"][c]) / r64["E"][c])) for c in range(len(specs)))
        corr = min(float(O.pc_correlations(u64[c], us[c]).min()) for c in range(len(specs)))
        errs[f] = {"eig_rel_err_max": rel, "pc_corr_min": corr}
        errs[f]["meets_bar"] = bar_met(errs[f])
    out["error_vs_fp64"] = errs
    if args.sequential:
        # the same cohorts one after another through the single-matrix path (generation untimed)
        t_seq = 0.0
        p = api.default_params(precision=forms[0], **kw)
        for sp in specs[: args.sequential]:
            ctx.generate(sp)
            t0 = time.perf_counter()
            ctx.run(p)
            t_seq += time.perf_counter() - t0
        scale = len(specs) / min(args.sequential, len(specs))
        out["sequential_single_path"] = {"form": forms[0], "cohorts_timed": min(args.sequential, len(specs)),
                                         "seconds_scaled_to_all": round(t_seq * scale, 3)}
    print(json.dumps(out), flush=True)
    return 0


def block_plan(p, nsamples, nsnps):
    """The library's block plan for `-m` (blocks of SNPs per pass)."""
    import ctypes as C
    from pcaone_amd import _lib
    bs, nb, bf = C.c_uint64(), C.c_uint64(), C.c_uint64()
    start = np.zeros(1 << 16, dtype=np.uint64)
    rc = _lib.lib().pcaone_block_plan(C.byref(p), nsamples, nsnps, C.byref(bs), C.byref(nb), C.byref(bf),
                                      start.ctypes.data, start.size)
    return {"blocksize": bs.value, "nblocks": nb.value, "band_factor": bf.value} if rc == 0 else None


def config_line(args) -> int:
    """One line of timings and errors (against the FP64 form) for another
    configuration, or for any of them through the out-of-core schedule
    (`--memory`, the reference's `-m`: blocks of SNPs per pass, genotypes
    resident in HBM)."""
    from pcaone_amd import api, synth
    if args.config == 4:
        return batch_line(args)
    run = dict(CONFIG_RUNS[args.config])
    run["params"] = dict(run["params"], **({"memory": args.memory} if args.memory > 0 else {}))
    spec = synth.config(args.config)
    ctx = api.Context(0)
    t0 = time.perf_counter()
    packed = None
    if args.stream:
        # the matrix stays in host memory; each product loads its -m block (data beyond HBM)
        require_mem = args.memory > 0
        if not require_mem:
            raise SystemExit("--stream needs --memory GB (the -m plan sets the blocks)")
        packed = synth.packed(spec)
        plan = block_plan(api.default_params(**run["params"]), spec.nsamples, spec.nsnps)
        log(f"config {args.config}: generated {spec.nsamples} x {spec.nsnps} on the host "
            f"({packed.nbytes / 1e9:.2f} GB) in {time.perf_counter() - t0:.1f}s")
        t0 = time.perf_counter()
        ctx.stream_host(packed, spec.nsamples, plan["blocksize"])
        log(f"config {args.config}: streamed open (allele counts over all blocks) {time.perf_counter() - t0:.1f}s")
    else:
        ctx.generate(spec)
        log(f"config {args.config}: generated {spec.nsamples} x {spec.nsnps} in {time.perf_counter() - t0:.1f}s")
    res, secs = {}, {}
    forms = [f for f in (args.forms.split(",") if args.forms else run["forms"]) if f != "fp64"] + ["fp64"]
    for form in forms:
        p = api.default_params(precision=form, verbose=int(args.verbose), **run["params"])
        if form != "fp64" and args.warmup:
            ctx.run(p)  # warm-up
        t0 = time.perf_counter()
        res[form] = ctx.run(p)
        secs[form] = time.perf_counter() - t0
        log(f"config {args.config} {form}: {secs[form]:.3f}s, epochs {res[form].info['total_epochs']}, "
            f"em {res[form].info['em_iters']}")
    ref = res["fp64"]
    out = {"config": args.config, "workload": f"{spec.nsamples} samples x {spec.nsnps} SNPs, {run['what']}",
           "seconds": {f: round(v, 4) for f, v in secs.items()},
           "epochs": {f: r.info["total_epochs"] for f, r in res.items()},
           "error_vs_fp64": {f: dict(errors(ref, r), meets_bar=bar_met(errors(ref, r))) for f, r in res.items()
                             if f != "fp64"},
           "g_ms_h_ms_per_run": None}
    if args.memory > 0:
        out["memory_gb"] = args.memory
        out["block_plan"] = block_plan(api.default_params(**run["params"]), spec.nsamples, spec.nsnps)
    if args.stream:
        bps = (spec.nsamples + 3) // 4
        blk = out["block_plan"]["blocksize"]
        out["streamed"] = {
            "host_matrix_bytes": int(packed.nbytes),
            # the two packed layouts of one block on the device, against both layouts of the whole matrix
            "device_genotype_bytes": int(2 * blk * bps),
            "resident_genotype_bytes": int(2 * spec.nsnps * bps),
            # per run of each form: blocks loaded, the host wall time of the loads and their rate
            # (every pass loads all blocks, spec.nsnps rows of bps bytes)
            "block_loads": {f: int(r.info["stream_loads"]) for f, r in res.items()},
            "load_s": {f: round(r.info["stream_load_ms"] / 1e3, 3) for f, r in res.items()},
            "load_gb_per_s": {f: round(r.info["stream_loads"] / out["block_plan"]["nblocks"] * spec.nsnps * bps
                                       / max(r.info["stream_load_ms"], 1e-9) / 1e6, 1) for f, r in res.items()},
            "what": "packed rows in host memory; every product loads its block (gather, DMA, interleave, "
                    "transpose, missing lists) into block-sized layouts"}
    ctx.time_kernels(True)
    rk = ctx.run(api.default_params(precision=forms[0], **run["params"]))
    ctx.time_kernels(False)
    out["g_ms_h_ms_per_run"] = [round(rk.info["g_ms"], 2), round(rk.info["h_ms"], 2),
                                rk.info["g_launches"], rk.info["h_launches"]]
    print(json.dumps(out), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nsamples", type=int, default=50_000)
    ap.add_argument("--nsnps", type=int, default=1_000_000)
    ap.add_argument("--k", type=int, default=40)
    ap.add_argument("--oversamples", type=int, default=20)
    ap.add_argument("--maxp", type=int, default=7)
    ap.add_argument("--cpu-snps", type=int, default=10_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--only", default="", help="comma list of forms to run (profiling)")
    ap.add_argument("--wm-g", type=int, default=0, help="GEMM wave tile of the G product (0 = tuned default)")
    ap.add_argument("--wm-h", type=int, default=0, help="GEMM wave tile of the H product (0 = tuned default)")
    ap.add_argument("--waves-g", type=int, default=0, help="waves per workgroup of the G product (0 = default)")
    ap.add_argument("--waves-h", type=int, default=0, help="waves per workgroup of the H product (0 = default)")
    ap.add_argument("--head", default="fp8", choices=sorted(DTYPE),
                    help="headline form (default: fp8, three E4M3 pieces; bf16 = hi/lo)")
    ap.add_argument("--config", type=int, default=3,
                    help="3: the headline workload; 2 (winSVD), 4 (batched cohorts) or 5 (EMU): a timing/error "
                         "line for that configuration at full size instead")
    ap.add_argument("--forms", default="", help="with --config: comma list of forms (default: all, fp64 last)")
    ap.add_argument("--cohorts", type=int, default=256, help="with --config 4: number of cohorts")
    ap.add_argument("--sequential", type=int, default=256,
                    help="with --config 4: cohorts also run one by one through the single-matrix path (0 = none)")
    ap.add_argument("--verbose", action="store_true", help="with --config: the library prints epochs and EM steps")
    ap.add_argument("--conv-maxp", type=int, default=40,
                    help="power iterations of the converged FP64 reference (0 = none)")
    ap.add_argument("--cpu-memory", type=float, default=2.0,
                    help="-m GB of the reference binary's CPU baseline (out-of-core mode)")
    ap.add_argument("--stream", action="store_true",
                    help="with --config and --memory: keep the matrix in host memory, stream the -m blocks")
    ap.add_argument("--memory", type=float, default=0.0,
                    help="with --config: run the out-of-core schedule (-m GB: blocks of SNPs per pass)")
    ap.add_argument("--copy-probe", action="store_true", help="run a 1 GiB plain copy (counter calibration)")
    ap.add_argument("--pcie", action="store_true",
                    help="also time the host-buffer boundary: .bed rows from host memory into HBM (upload + finalize)")
    ap.add_argument("--kchunk-g", type=int, default=0, help="split-K chunk of the G product in codes (0 = auto)")
    ap.add_argument("--kchunk-h", type=int, default=0, help="split-K chunk of the H product in codes (0 = auto)")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "one GPU only: the randomized-SVD hot path is measured on a single MI355X; "
                                   "multi-GPU sharding is not run",
                          "n_gpus": args.gpus}))
        return 2

    from pcaone_amd import api, synth

    if args.config != 3 or args.memory > 0:
        return config_line(args)
    spec = synth.config(3, nsamples=args.nsamples, nsnps=args.nsnps)
    k, os_, maxp = args.k, args.oversamples, args.maxp
    L = k + max(os_, k)  # src/Cmd.cpp:216
    ctx = api.Context(0)
    t0 = time.perf_counter()
    ctx.generate(spec)
    setup_s = time.perf_counter() - t0
    log(f"generated {spec.nsamples} x {spec.nsnps} genotypes in HBM in {setup_s:.1f}s")

    def params(form):
        return api.default_params(svd=1, k=k, oversamples=os_, maxp=maxp, tol=0.0, precision=form,
                                  wm_g=args.wm_g, wm_h=args.wm_h, waves_g=args.waves_g, waves_h=args.waves_h,
                                  kchunk_g=args.kchunk_g, kchunk_h=args.kchunk_h)

    if args.copy_probe:
        ctx.copy_probe(1 << 30)
        log("copy probe: 1 GiB")
    only = set(args.only.split(",")) if args.only else None
    results = {}
    omega_ms = {}  # host time of the Omega draw inside each timed run, per form

    def timed(form, steps, warmup):
        for i in range(warmup):
            r = ctx.run(params(form))
            log(f"{form} warmup {i}: {r.info['seconds']:.3f}s")
        ts = []
        r = None
        omega_ms.setdefault(form, [])
        for i in range(steps):
            t0 = time.perf_counter()
            r = ctx.run(params(form))
            ts.append(time.perf_counter() - t0)
            omega_ms[form].append(r.info["omega_ms"])
            log(f"{form} step {i}: {ts[-1]:.3f}s ({r.info['epochs']} epochs, Omega draw {r.info['omega_ms']:.1f} ms)")
        return r, ts

    head = args.head
    r_head, ts_head = (None, [])
    if only is None or head in only:
        r_head, ts_head = timed(head, args.steps, args.warmup)
        results[head] = r_head
        # live kernel timing (HIP events on the context stream), separate run
        ctx.time_kernels(True)
        rk = ctx.run(params(head))
        ctx.time_kernels(False)
        kt = rk.info
    if only is None or "fp32" in only:
        r_fp32, ts_fp32 = timed("fp32", 1, 1 if args.warmup else 0)
        results["fp32"] = r_fp32
    ts_side = {}
    side_kt = {}
    for side in SIDE_FORMS:
        if only is None and head != side:
            # reported beside the headline with its errors and its product launch times
            results[side], ts_side[side] = timed(side, 1, 1 if args.warmup else 0)
            ctx.time_kernels(True)
            side_kt[side] = ctx.run(params(side)).info
            ctx.time_kernels(False)
    if (only is None and not args.no_fp64) or (only and "fp64" in only):
        r64, ts64 = timed("fp64", 1, 0)
        results["fp64"] = r64
    if only is not None and head not in only:
        print(json.dumps({f: {"seconds": results[f].info["seconds"]} for f in results}))
        return 0
    if only is not None:
        print(json.dumps({"form": head, "seconds": float(np.mean(ts_head)), "g_ms": kt["g_ms"] / kt["g_launches"],
                          "h_ms": kt["h_ms"] / kt["h_launches"], "wm_g": args.wm_g, "wm_h": args.wm_h,
                          "waves_g": args.waves_g, "waves_h": args.waves_h,
                          "kchunk_g": args.kchunk_g, "kchunk_h": args.kchunk_h}), flush=True)
        return 0

    N, M = spec.nsamples, spec.nsnps
    flops_launch = 2.0 * N * M * L  # one product over the whole matrix, algorithmic
    g_avg = kt["g_ms"] / max(kt["g_launches"], 1)
    h_avg = kt["h_ms"] / max(kt["h_launches"], 1)
    dom, dom_ms = ("H = A^T G", h_avg) if h_avg >= g_avg else ("G = A Omega", g_avg)
    achieved = flops_launch / (dom_ms * 1e-3) / 1e12
    traffic = recorded_traffic(N, M, L, head, "H" if dom.startswith("H") else "G")
    value = float(np.mean(ts_head))
    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(value * 1e3, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE[head],
        "data": "synthetic Balding-Nichols genotypes generated in HBM (no dataset download)",
        "config": {"workload": f"precision form: {N} samples x {M} SNPs, 8 populations, --svd 1, k={k}, "
                               f"oversamples {os_}->{L - k}, L={L}, {maxp} power iterations ({maxp + 1} passes)",
                   "nsamples": N, "nsnps": M, "k": k, "L": L, "passes": maxp + 1,
                   "form": FORM_TEXT.get(head, head)},
        "roofline": {"bound": "mfma", "kernel": f"packed_gemm ({dom})", "achieved": round(achieved, 2),
                     "peak": PEAK_TFLOPS[head], "unit": "TOP/s" if head == "int8" else "TFLOP/s",
                     "frac": round(achieved / PEAK_TFLOPS[head], 4),
                     "traffic": traffic, "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE)",
                     "avg_launch_ms": round(dom_ms, 3),
                     "algorithmic_flop_per_launch": flops_launch,
                     "executed_frac": round(PIECES[head] * achieved / PEAK_TFLOPS[head], 4),
                     "other_product_avg_ms": round(g_avg if dom.startswith("H") else h_avg, 3)},
        "setup_s": round(setup_s, 2),
        # every step draws Omega again on the host (libstdc++ engine, parallel), as the reference's
        # computeUSV does at epoch 0; this is that draw's share of ms_per_step
        "omega_draw_ms_per_step": round(float(np.mean(omega_ms[head])), 2),
    }
    if "fp32" in results:
        t32 = float(np.mean(ts_fp32))
        out["fp32_form"] = {"seconds": round(t32, 4), f"speedup_{head}_over_fp32": round(t32 / value, 2)}
        for side, ts in ts_side.items():
            kt_s = side_kt[side]
            g_s = kt_s["g_ms"] / max(kt_s["g_launches"], 1)
            out[f"{side}_form"] = {"seconds": round(float(np.mean(ts)), 4),
                                   f"speedup_{side}_over_fp32": round(t32 / float(np.mean(ts)), 2),
                                   "g_launch_ms": round(g_s, 3),
                                   "h_launch_ms": round(kt_s["h_ms"] / max(kt_s["h_launches"], 1), 3),
                                   "algorithmic_tflops": round(flops_launch / (g_s * 1e-3) / 1e12, 1),
                                   "peak": PEAK_TFLOPS[side],
                                   "frac": round(flops_launch / (g_s * 1e-3) / 1e12 / PEAK_TFLOPS[side], 4),
                                   "form": FORM_TEXT[side]}
    if "fp64" in results:
        ref = results["fp64"]
        out["fp64_form_seconds"] = round(ts64[0], 3)
        out["error_vs_fp64"] = {f: errors(ref, results[f]) for f in results if f != "fp64"}
        # the bar each form is held to against the FP64 form of the same run
        for f, e in out["error_vs_fp64"].items():
            e["meets_bar"] = bar_met(e)
        if args.conv_maxp > 0:
            # The same workload in FP64 iterated until diff < 1e-12 (or conv_maxp power
            # iterations): the distance of every 8-pass form from the converged top-k.
            # The noise PCs beyond the npops - 1 structured ones have gaps of 1e-4
            # and converge slowly, so correlations are also given for the structured PCs.
            t0 = time.perf_counter()
            rc = ctx.run(api.default_params(svd=1, k=k, oversamples=os_, maxp=args.conv_maxp, tol=1e-12,
                                            precision="fp64"))
            t_conv = time.perf_counter() - t0
            ns = spec.npops - 1
            conv = {}
            for f in results:
                e = errors(rc, results[f])
                from oracle import pcaone_oracle as O  # checker only
                e["pc_corr_min_structured"] = float(O.pc_correlations(rc.U[:, :ns], results[f].U[:, :ns]).min())
                e["eig_rel_err_max_structured"] = float(np.max(np.abs(results[f].E[:ns] - rc.E[:ns]) / rc.E[:ns]))
                conv[f] = e
            out["error_vs_converged_fp64"] = conv
            out["converged_fp64"] = {"epochs": rc.info["epochs"], "last_diff": rc.trace[-1] if rc.trace else None,
                                     "seconds": round(t_conv, 2), "structured_pcs": ns}
            log(f"converged FP64 reference: {rc.info['epochs']} epochs in {t_conv:.1f}s")
    if args.pcie:
        packed = ctx.download_packed()
        t0 = time.perf_counter()
        ctx.upload_packed(packed, N)
        t_up = time.perf_counter() - t0
        out["pcie_inclusive"] = {"upload_finalize_s": round(t_up, 3),
                                 "host_GB_per_s": round(packed.nbytes / t_up / 1e9, 2),
                                 "time_to_k_pcs_from_host_s": round(t_up + value, 3)}
        del packed
    if not args.no_cpu:
        try:
            log("cpu baseline ...")
            out["cpu_baseline"] = cpu_baseline(args, spec, k, os_, maxp)
        except Exception as e:  # reported, never silently replaced
            out["cpu_baseline"] = {"error": str(e)}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
