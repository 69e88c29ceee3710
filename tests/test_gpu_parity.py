"""GPU parity of the HIP path against the FP64 oracle and the reference's golden outputs.

Tolerances: the FP64 form must reproduce the reference outputs to their print
precision (6 significant digits); the FP32 and BF16 hi/lo forms must meet the
accuracy bar against the FP64 oracle on the same Omega: eigenvalue relative
error <= 1e-4 and every PC correlated >= 0.9999; so must the int8 form (two
int8 pieces of 16-bit per-column fixed point, exact int32 accumulation) and
the fp8 form (three E4M3 pieces per value, FP32 accumulation).  The
single-piece forms (FP16, BF16MIX, FP16MIX) meet the eigenvalue bar and
correlate >= 0.999.
"""
from pathlib import Path

import numpy as np
import pytest

from conftest import GOLDEN, golden_cases, load_meta, params_from_args
from oracle import pcaone_oracle as O
from pcaone_amd import api

pytestmark = pytest.mark.gpu

DMAP = np.array([2.0, 0.0, 0.0, -2.0])


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def ctx():
    c = api.Context(0)
    yield c
    c.close()


def pack(codes: np.ndarray) -> np.ndarray:
    m, n = codes.shape
    bps = (n + 3) // 4
    out = np.zeros((m, bps), dtype=np.uint8)
    for e in range(4):
        sub = codes[:, e::4].astype(np.uint8)
        out[:, : sub.shape[1]] |= sub << (2 * e)
    return out


# ---------------------------------------------------------------- kernel level
@pytest.mark.parametrize("form", ["fp64", "fp32", "bf16", "bf16x1", "fp16"])
@pytest.mark.parametrize("nrows,ncodes,L", [(1, 1, 1), (37, 127, 20), (300, 129, 80), (513, 1000, 128),
                                            (64, 4096, 16), (257, 333, 40)])
def test_packed_gemm_exact(ctx, form, nrows, ncodes, L):
    say(f"packed_gemm form={form} nrows={nrows} ncodes={ncodes} L={L}")
    rng = np.random.default_rng(nrows * 7 + ncodes + L)
    codes = rng.integers(0, 4, (nrows, ncodes)).astype(np.uint8)
    W = rng.integers(-64, 64, (ncodes, L)) / 8.0  # exact in every operand type
    Y = ctx.packed_gemm(form, pack(codes), ncodes, W)
    np.testing.assert_array_equal(Y, DMAP[codes] @ W)


@pytest.mark.parametrize("nrows,ncodes,L", [(1, 1, 1), (37, 127, 20), (300, 129, 80), (513, 1000, 128),
                                            (64, 4096, 16), (257, 333, 40), (100, 65536, 24)])
def test_packed_gemm_int8_exact(ctx, nrows, ncodes, L):
    """W = q 2^-7 with |q| <= 32639 and each column's maximum exactly 32639 2^-7:
    the column scale is 2^-7, the fixed point is exact and so is the product."""
    say(f"packed_gemm form=int8 nrows={nrows} ncodes={ncodes} L={L}")
    rng = np.random.default_rng(nrows * 5 + ncodes + L)
    codes = rng.integers(0, 4, (nrows, ncodes)).astype(np.uint8)
    q = rng.integers(-32639, 32640, (ncodes, L))
    q[rng.integers(0, ncodes, L), np.arange(L)] = rng.choice([-32639, 32639], L)
    W = q / 128.0
    Y = ctx.packed_gemm("int8", pack(codes), ncodes, W)
    np.testing.assert_array_equal(Y, DMAP[codes] @ W)


@pytest.mark.parametrize("nrows,ncodes,L", [(1, 1, 1), (37, 127, 20), (300, 129, 80), (513, 1000, 128),
                                            (257, 333, 40), (64, 1024, 16)])
def test_packed_gemm_fp8_exact(ctx, nrows, ncodes, L):
    """W = q 2^-7 with |q| <= 4095 and each column's maximum exactly 4095 2^-7:
    the column scale is 2^-3, so W / s = q / 16 has at most 12 significant bits
    above 2^-4 and three E4M3 pieces (4 bits each, round to nearest) hold it
    exactly; every partial sum is a multiple of 2^-3 below 2^21, exact in FP32."""
    say(f"packed_gemm form=fp8 nrows={nrows} ncodes={ncodes} L={L}")
    rng = np.random.default_rng(nrows * 3 + ncodes + L)
    codes = rng.integers(0, 4, (nrows, ncodes)).astype(np.uint8)
    q = rng.integers(-4095, 4096, (ncodes, L))
    q[rng.integers(0, ncodes, L), np.arange(L)] = rng.choice([-4095, 4095], L)
    W = q / 128.0
    Y = ctx.packed_gemm("fp8", pack(codes), ncodes, W)
    np.testing.assert_array_equal(Y, DMAP[codes] @ W)


def test_packed_gemm_fp8_small_and_signed_values(ctx):
    """Columns of one value each: a power of two, a value needing all three
    pieces, values near the top of E4M3 after scaling, and a zero column."""
    say("packed_gemm fp8 special columns")
    codes = np.array([[0, 3, 2, 0, 1, 3, 0, 0]], dtype=np.uint8)
    n = codes.shape[1]
    W = np.zeros((n, 5))
    W[:, 0] = 2.0 ** -20
    W[:, 1] = 1.0 + 2.0 ** -11          # 12 significant bits
    W[:, 2] = -(447.0 + 0.5)            # scaled into (224, 448]
    W[:, 3] = 3.0 * 2.0 ** 20
    Y = ctx.packed_gemm("fp8", pack(codes), n, W)
    np.testing.assert_array_equal(Y, DMAP[codes] @ W)


def test_packed_gemm_rounding_fp32(ctx):
    say("packed_gemm fp32 rounding")
    rng = np.random.default_rng(11)
    codes = rng.integers(0, 4, (200, 3000)).astype(np.uint8)
    W = rng.standard_normal((3000, 24))
    ref = DMAP[codes] @ W
    for form, tol in (("fp64", 1e-12), ("fp32", 2e-5), ("bf16", 3e-5), ("bf16x1", 2e-2), ("fp16", 3e-3),
                      ("int8", 2e-4), ("fp8", 3e-4)):
        Y = ctx.packed_gemm(form, pack(codes), 3000, W)
        err = np.abs(Y - ref).max() / np.abs(ref).max()
        say(f"  {form}: max rel err {err:.2e}")
        assert err < tol


# ---------------------------------------------------------------- data level
def test_allele_freq_bit_exact(ctx, golden_inputs):
    for name in ("f1", "f4"):
        say(f"allele_freq {name}")
        packed, n, m = O.read_plink(golden_inputs[name])
        ctx.read_bed(golden_inputs[name])
        F = ctx.allele_freq()
        Fo = O.allele_freq(O.decode_codes(packed, n))
        assert np.array_equal(F, Fo)
        _, _, miss = ctx.dims()
        assert miss == int((O.decode_codes(packed, n) == 1).sum())


def test_allele_freq_ragged_and_edge_rows(ctx):
    say("allele_freq ragged")
    rng = np.random.default_rng(3)
    for n in (1, 2, 3, 5, 129, 1000):
        codes = rng.integers(0, 4, (70, n)).astype(np.uint8)
        codes[0] = 1  # all missing -> F = 0
        codes[1] = 3  # monomorphic
        ctx.upload_packed(pack(codes), n)
        assert np.array_equal(ctx.allele_freq(), O.allele_freq(codes))


# ---------------------------------------------------------------- operator steps
def test_operator_epoch_matches_oracle(ctx, golden_inputs):
    say("operator steps fp64 f1")
    prefix = golden_inputs["f1"]
    packed, n, m = O.read_plink(prefix)
    k, L = 5, 15
    ctx.read_bed(prefix)
    ctx.configure("fp64", L)
    ctx.standardize(2)
    om = O.omega(112, True, n, L)
    ctx.set_omega(om)
    ctx.sketch(0, m, 0)
    d = O.Data.from_packed(packed, n, O.Params(svd=1, k=k))
    d.standardize_E()
    G = d.G.T @ om
    H = d.G @ G
    # the products themselves, as a caller keeping a host computeUSV reads them
    np.testing.assert_allclose(ctx.get_g(0, m, L), G, rtol=1e-11, atol=1e-11 * np.abs(G).max())
    np.testing.assert_allclose(ctx.get_g(100, 357, L), G[100:357], rtol=1e-11, atol=1e-11 * np.abs(G).max())
    np.testing.assert_allclose(ctx.get_h(0, L), H, rtol=1e-11, atol=1e-11 * np.abs(H).max())
    sig, _ = ctx.finish_epoch(k, L, False)
    Q1, R1 = np.linalg.qr(G)
    Q2, R2 = np.linalg.qr(Q1)
    B = np.linalg.solve((R2 @ R1).T, H.T)
    _, s, vt = np.linalg.svd(B, full_matrices=False)
    np.testing.assert_allclose(sig, s, rtol=1e-11)
    U = ctx.get_u(k)
    np.testing.assert_allclose(O.sign_align(vt.T[:, :k], U), vt.T[:, :k], atol=1e-10)
    ctx.save_prev()
    # second epoch: Omega <- thin-Q(H) with flipOmg, then the products again
    ctx.update_omega(0)
    ctx.sketch(0, m, 0)
    sig2, diff = ctx.finish_epoch(k, L, True)
    om2 = O.flip_omg(om, O.thin_q(H))
    G = d.G.T @ om2
    H = d.G @ G
    Q1, R1 = np.linalg.qr(G)
    Q2, R2 = np.linalg.qr(Q1)
    B = np.linalg.solve((R2 @ R1).T, H.T)
    _, s2, vt2 = np.linalg.svd(B, full_matrices=False)
    np.testing.assert_allclose(sig2, s2, rtol=1e-11)
    assert abs(diff - (1 - O.mev(vt2.T[:, :k], vt.T[:, :k]))) < 1e-11


def test_operator_thin_q_without_flip(ctx, golden_inputs):
    """The out-of-core --svd 1 branch's Omega update through the C ABI:
    thin-Q of H with Householder signs and no flipOmg, then the products."""
    say("operator thin-Q without flipOmg")
    prefix = golden_inputs["f1"]
    packed, n, m = O.read_plink(prefix)
    k, L = 5, 15
    ctx.read_bed(prefix)
    ctx.configure("fp64", L)
    ctx.standardize(2)
    om = O.omega(112, True, n, L)
    ctx.set_omega(om)
    ctx.sketch(0, m, 0)
    d = O.Data.from_packed(packed, n, O.Params(svd=1, k=k))
    d.standardize_E()
    H = d.G @ (d.G.T @ om)
    ctx.thin_q_omega(0)
    ctx.sketch(0, m, 0)
    G2 = d.G.T @ O.thin_q(H)
    np.testing.assert_allclose(ctx.get_g(0, m, L), G2, rtol=1e-10, atol=1e-10 * np.abs(G2).max())


@pytest.mark.parametrize("emu", [False, True])
def test_missing_entries_with_chunked_sources(ctx, emu):
    """Missing genotypes at a size where the FP64 form splits the H side's
    compact source rows into two L2-sized index chunks (partial rows summed in
    chunk order): the run matches the FP64 oracle."""
    from pcaone_amd import synth
    spec = synth.Spec(seed=77, nsamples=200, nsnps=50000, npops=3, miss_rate=0.1, high_miss_frac=0.05,
                      high_miss_rate=0.3)
    say(f"missing entries, chunked sources, emu={emu}")
    ctx.generate(spec)
    kw = dict(svd=1, k=4, oversamples=4, maxp=3, tol=0.0, emu=emu, maxiter=2, tolem=0.0)
    ref = O.run_pca_with_halko(O.Data.from_packed(synth.packed(spec), spec.nsamples, O.Params(**kw)))
    kw["emu"] = int(emu)
    r = ctx.run(api.default_params(precision="fp64", **kw))
    np.testing.assert_allclose(r.S, ref.S, rtol=1e-9)
    np.testing.assert_allclose(O.sign_align(ref.U, r.U), ref.U, atol=1e-8)


TILE_SPECS = {  # 1,000 x 3,000: 16 row blocks of SNPs, 6 of samples
    # dense: tiles bounded by the entry cap (about 30 indexes each), several per row block
    "dense": dict(seed=31, nsamples=1000, nsnps=3000, npops=3, miss_rate=0.5, high_miss_frac=0.05,
                  high_miss_rate=0.9),
    # sparse: tiles bounded by the index span the source rows may take in LDS
    "sparse": dict(seed=32, nsamples=1000, nsnps=3000, npops=3, miss_rate=0.02),
}


@pytest.mark.parametrize("density", ["dense", "sparse"])
@pytest.mark.parametrize("L", [4, 12, 20, 32])
def test_tiled_corrections_match_oracle(ctx, density, L):
    """The reduced forms' missing-entry corrections run on tiles of the two
    lists (row blocks of 192 rows; source rows, indexes and FP32 weights
    staged in LDS).  L = 4, 12, 20, 32 take 1, 3, 5 and 8 lanes per entry.
    The products of whole and cut windows (a window starting inside a row
    block, tiles cut by the window's SNP range, several parts per row block)
    equal A Omega and A^T G of the FP64 oracle, where missing entries are 0
    after centering."""
    from pcaone_amd import synth
    spec = synth.Spec(**TILE_SPECS[density])
    say(f"tiled corrections {density} L={L}")
    ctx.generate(spec)
    n, m = spec.nsamples, spec.nsnps
    d = O.Data.from_packed(synth.packed(spec), n, O.Params(svd=1, k=2))
    d.standardize_E()
    om = O.omega(112, True, n, L)
    ctx.configure("fp32", L)
    ctx.standardize(2)  # scales the columns once: sc *= sqrt(2) / sd
    for form, tol in (("fp32", 1e-5), ("bf16", 1e-4)):
        ctx.configure(form, L)
        ctx.set_omega(om)
        for j0, j1 in ((0, m), (100, 357), (383, 1900)):
            ctx.sketch(j0, j1, 0)
            G = d.G[:, j0:j1].T @ om
            H = d.G[:, j0:j1] @ G
            np.testing.assert_allclose(ctx.get_g(j0, j1, L), G, rtol=0, atol=tol * np.abs(G).max())
            np.testing.assert_allclose(ctx.get_h(0, L), H, rtol=0, atol=tol * np.abs(H).max())


@pytest.mark.parametrize("svd", [1, 2])
def test_tiled_emu_runs_match_gather_and_oracle(ctx, svd):
    """EMU runs with the tiled corrections and tiled imputation weights (U S
    and V rows staged in LDS) against the per-entry gather form of the same
    precision and against the FP64 oracle."""
    from pcaone_amd import synth
    spec = synth.Spec(seed=33, nsamples=1000, nsnps=3000, npops=3, miss_rate=0.2, high_miss_frac=0.05,
                      high_miss_rate=0.6)
    say(f"tiled EMU run svd={svd}")
    ctx.generate(spec)
    kw = dict(svd=svd, k=4, oversamples=6, maxp=4, tol=0.0, bands=4, emu=True, maxiter=3, tolem=0.0)
    ref = O.run_pca_with_halko(O.Data.from_packed(synth.packed(spec), spec.nsamples, O.Params(**kw)))
    kw["emu"] = 1
    tiled = ctx.run(api.default_params(precision="fp32", **kw))
    gather = ctx.run(api.default_params(precision="fp32", miss_gather=1, **kw))
    assert tiled.info["em_iters"] == gather.info["em_iters"] == len(ref.em_trace)
    np.testing.assert_allclose(tiled.E, gather.E, rtol=1e-6)
    assert O.pc_correlations(gather.U, tiled.U).min() > 1 - 1e-6
    assert np.max(np.abs(tiled.E - ref.E) / ref.E) <= 1e-4
    assert O.pc_correlations(ref.U, tiled.U).min() >= 0.9999


@pytest.mark.parametrize("k", [8, 20, 30, 40, 50])
def test_thin_q_paths_match_oracle(ctx, golden_inputs, k):
    """The Omega update (thin-Q of H with Householder signs, then flipOmg) runs
    its L x L steps on the device up to L = 80 (k = 8, 20, 30, 40: one to five
    16 x 16 register blocks per thread) and on the host above (k = 50: L = 100);
    every path matches the FP64 oracle over 5 epochs."""
    say(f"thin-Q path k={k}")
    prefix = golden_inputs["cfg1"]
    packed, n, m = O.read_plink(prefix)
    kw = dict(svd=1, k=k, oversamples=k, maxp=4, tol=0.0)
    ref = O.run_pca_with_halko(O.Data.from_packed(packed, n, O.Params(**kw)))
    ctx.read_bed(prefix)
    r = ctx.run(api.default_params(precision="fp64", **kw))
    np.testing.assert_allclose(r.S, ref.S, rtol=1e-9)
    np.testing.assert_allclose(O.sign_align(ref.U, r.U), ref.U, atol=1e-7)


# ---------------------------------------------------------------- whole runs vs golden
def run_case(ctx, golden_inputs, case, form, golden_pgen=None):
    """One golden case on the device: the .bed read, or for a --pgen case the
    .pgen the reference read (its -m runs then take the .pgen shuffle)."""
    meta = load_meta(case)
    params, printv = params_from_args(meta["args"], pgen="pgen" in meta)
    say(f"golden {case} form={form} args='{meta['args']}'")
    if "pgen" in meta:
        ctx.read_pgen(golden_pgen(case), maf=params.maf, hardcall=params.hardcall)
    else:
        ctx.read_bed(golden_inputs[meta["input"]], maf=params.maf)
    p = api.default_params(svd=params.svd, k=params.k, oversamples=params.oversamples, maxp=params.maxp,
                           tol=params.tol, seed=params.seed, rand=params.rand, bands=params.bands,
                           noshuffle=int(params.noshuffle), emu=int(params.emu), maxiter=params.maxiter,
                           tolem=params.tolem, memory=params.memory, buffer=params.buffer, ld=int(params.ld),
                           ld_stats=params.ld_stats, threads=params.threads, hardcall=int(params.hardcall),
                           precision=form)
    return meta, params, printv, ctx.run(p, want_v=printv)


@pytest.mark.parametrize("case", golden_cases())
def test_golden_fp64(ctx, golden_inputs, golden_pgen, case):
    meta, params, printv, r = run_case(ctx, golden_inputs, case, "fp64", golden_pgen)
    d = GOLDEN / case
    if meta["trace"]:
        np.testing.assert_allclose(r.trace, meta["trace"], atol=1.5e-6)
    if meta["stop_epochs"]:
        assert r.info["epochs"] == meta["stop_epochs"][-1]
    if meta["em_trace"]:
        np.testing.assert_allclose(r.em_trace, meta["em_trace"], atol=1.5e-6)
    if (d / "eigvals").stat().st_size > 1:
        np.testing.assert_allclose(r.E, O.read_matrix(d / "eigvals").ravel(), rtol=1e-5)
        np.testing.assert_allclose(r.S, O.read_matrix(d / "sigvals").ravel(), rtol=1e-5)
    Uref = O.read_matrix(d / "eigvecs")
    np.testing.assert_allclose(O.sign_align(Uref, r.U), Uref, atol=2e-6)
    if params.svd == 3:  # exact PCA: flip_UV fixes the signs (src/Main.cpp:215)
        np.testing.assert_allclose(r.U, Uref, atol=2e-6)
    if printv and (d / "loadings").exists():
        Vref = O.read_matrix(d / "loadings")
        np.testing.assert_allclose(O.sign_align(Vref, r.V), Vref, atol=2e-6)


REDUCED = [(c, f) for f in ("fp32", "bf16", "fp16", "int8", "fp8")
           for c in ("f1_d1", "f1_d2", "f4_d1_miss", "f4_emu_d1", "cfg1_d1", "cfg1_d2")]
# The mixed form (one BF16 piece of Omega) meets the bar on --svd 1 only; on
# winSVD the rounded per-window Omega moves the unconverged noise PCs.
REDUCED += [(c, f) for f in ("bf16mix", "fp16mix") for c in ("f1_d1", "f4_d1_miss", "cfg1_d1")]
# degenerate SNPs (monomorphic, all missing, one genotype among missing entries)
REDUCED += [(c, f) for f in ("fp32", "bf16") for c in ("f5_d1", "f5_d2", "f5_emu")]
# rank deficient (rank 11 < L = 16): the orthonormalization fallbacks in every form
REDUCED += [(c, f) for f in ("fp32", "bf16") for c in ("f6_d1", "f6_d2")]
# the out-of-core schedule (block windows, permute_plink order, per-pass imputation)
REDUCED += [(c, f) for f in ("fp32", "bf16") for c in ("f1_d2_ooc", "f4_emu_d2_ooc")]
# .pgen input (every record type; the .pgen -m shuffle)
REDUCED += [(c, f) for f in ("fp32", "bf16") for c in ("f10_pg_d1", "f1_pg_ooc")]
# Single-piece forms round Omega (or G) to 8-11 bits: the eigenvalues meet the
# bar, the unconverged trailing PCs move by up to ~1e-3 in correlation.
PC_CORR_MIN = {"fp32": 0.9999, "bf16": 0.9999, "int8": 0.9999, "fp8": 0.9999, "fp16": 0.999, "bf16mix": 0.999,
               "fp16mix": 0.999}


@pytest.mark.parametrize("case,form", REDUCED)
def test_golden_reduced_precision(ctx, golden_inputs, golden_pgen, case, form):
    meta, params, printv, r = run_case(ctx, golden_inputs, case, form, golden_pgen)
    ref = O.run_plink(golden_inputs[meta["input"]], params)
    rel = np.abs(r.E - ref.E) / ref.E
    corr = O.pc_correlations(ref.U, r.U)
    say(f"  eig rel err max {rel.max():.2e}  min PC corr {corr.min():.8f}")
    assert rel.max() <= 1e-4
    assert corr.min() >= PC_CORR_MIN[form]


@pytest.mark.parametrize("form", ["bf16", "fp32", "int8", "fp8"])
def test_precision_form_at_l80_matches_oracle(ctx, golden_inputs, form):
    """The precision workload's settings (k = 40, oversampling 20 raised to 40
    by src/Cmd.cpp:216, so L = 80; 7 power iterations, --tol-rsvd 0) on the
    2,000 x 20,000 input: the reduced forms (BF16 hi/lo with the fused G
    epilogue and the device thin-Q at L = 80, FP32, int8, fp8) against the FP64
    oracle on the same Omega.  Bar: eigenvalue relative error <= 1e-4 and
    every one of the 40 PCs correlated >= 0.9999."""
    say(f"precision form {form} at k=40, L=80 on cfg1")
    prefix = golden_inputs["cfg1"]
    kw = dict(svd=1, k=40, oversamples=20, maxp=7, tol=0.0)
    ref = O.run_plink(prefix, O.Params(**kw))
    ctx.read_bed(prefix)
    r = ctx.run(api.default_params(precision=form, **kw))
    assert r.info["epochs"] == 8 and r.U.shape == (2000, 40)
    rel = np.abs(r.E - ref.E) / ref.E
    corr = O.pc_correlations(ref.U, r.U)
    ang = O.principal_angles(ref.U, r.U)
    say(f"  eig rel err max {rel.max():.2e}  min PC corr {corr.min():.8f}  max angle {ang.max():.2e}")
    assert rel.max() <= 1e-4
    assert corr.min() >= 0.9999


@pytest.mark.parametrize("case,form", [("f6_d1", "fp64"), ("f6_d1", "bf16"), ("f6_d2", "fp64"),
                                       ("f6_near", "fp64")])
def test_rank_deficient_input_takes_the_fallbacks(ctx, golden_inputs, case, form):
    """12 distinct samples repeated 16 times: rank 11.  With L = 16 the Gram
    matrices of G and H are singular, where CholeskyQR2 breaks down; the thin Q
    of H falls back to a Householder QR and the QR of G to a rank-revealing
    second pass, and the run matches the reference binary (golden FP64 test)
    and the oracle.  With L = 10 (f6_near, one under the rank) no fallback is
    needed for H."""
    meta, params, printv, r = run_case(ctx, golden_inputs, case, form)
    info = r.info
    say(f"  {case} {form}: householder {info['householder_fallbacks']} second passes "
        f"{info['gram_second_passes']} epoch rollbacks {info['epoch_rollbacks']} epochs {info['epochs']}")
    ref = O.run_plink(golden_inputs[meta["input"]], params)
    rel = np.abs(r.E - ref.E) / ref.E
    corr = O.pc_correlations(ref.U, r.U)
    say(f"  eig rel err max {rel.max():.2e}  min PC corr {corr.min():.10f}")
    assert rel.max() <= (1e-10 if form == "fp64" else 1e-4)
    assert corr.min() >= (1 - 1e-9 if form == "fp64" else 0.9999)
    if case == "f6_near":
        assert info["householder_fallbacks"] == 0
    else:
        assert info["householder_fallbacks"] >= 1 and info["gram_second_passes"] >= 1
    # winSVD checks its Omega updates' thin-Q on the device and reruns an
    # epoch whose check failed with the host test (and its Householder QR)
    assert (info["epoch_rollbacks"] >= 1) == (case == "f6_d2")


def test_well_conditioned_runs_take_no_fallback(ctx, golden_inputs):
    meta, params, printv, r = run_case(ctx, golden_inputs, "cfg1_d1", "bf16")
    assert r.info["householder_fallbacks"] == 0 and r.info["gram_second_passes"] == 0
    meta, params, printv, r = run_case(ctx, golden_inputs, "cfg1_d2", "bf16")
    assert r.info["householder_fallbacks"] == 0 and r.info["epoch_rollbacks"] == 0


@pytest.mark.parametrize("form", ["fp64", "bf16"])
def test_winsvd_windows_past_the_last_snp(ctx, form):
    """M = 1,008 < bands^2 = 4,096 with -w 64: windows of ceil(1008 / 64) = 16
    SNPs, so window 63 starts at SNP 1,008 = M and is empty (src/Halko.cpp:
    191-194 give it width 0; the golden case f9_d2_empty holds the reference's
    run).  Every other window has 16 SNPs, more than L = 14, so no H of a
    band is rank deficient and the trajectory is fixed by the data (with a
    shorter last window H can lose rank, and the Householder completion of
    Omega then depends on rounding).  Here against the oracle."""
    from pcaone_amd import synth
    spec = synth.Spec(seed=7, nsamples=150, nsnps=1008, npops=3)
    say(f"winSVD with empty trailing windows, {form}")
    kw = dict(svd=2, k=4, oversamples=10, bands=64, maxp=7, tol=0.0)
    ref = O.run_pca_with_halko(O.Data.from_packed(synth.packed(spec), spec.nsamples, O.Params(**kw)))
    ctx.generate(spec)
    r = ctx.run(api.default_params(precision=form, **kw))
    assert r.info["epochs"] == ref.epochs == 8
    if form == "fp64":
        np.testing.assert_allclose(r.trace, ref.trace, atol=1e-9)
        np.testing.assert_allclose(r.S, ref.S, rtol=1e-9)
        np.testing.assert_allclose(O.sign_align(ref.U, r.U), ref.U, atol=1e-7)
    else:
        assert (np.abs(r.E - ref.E) / ref.E).max() <= 1e-4
        assert O.pc_correlations(ref.U, r.U).min() >= 0.9999


def test_bed_bad_magic_through_the_device_reader(ctx, tmp_path):
    """pcaone_bed_read checks the magic before any upload (src/FilePlink.hpp:24-27)."""
    say("bad .bed magic")
    prefix = tmp_path / "bad"
    (tmp_path / "bad.fam").write_text("a a 0 0 0 -9\nb b 0 0 0 -9\n")
    (tmp_path / "bad.bim").write_text("1 s1 0 1 A G\n")
    (tmp_path / "bad.bed").write_bytes(b"\x6c\x1b\x00\x00")
    with pytest.raises(api._lib.PcaoneError, match="Incorrect magic number in plink bed file.") as e:
        ctx.read_bed(str(prefix))
    assert e.value.status == 4


def test_bad_shapes_are_errors(ctx, golden_inputs):
    say("error paths")
    ctx.read_bed(golden_inputs["f1"])
    with pytest.raises(api._lib.PcaoneError) as e:
        ctx.run(api.default_params(svd=1, k=150, oversamples=150, maxp=1))
    assert e.value.status == 1
    with pytest.raises(api._lib.PcaoneError):
        ctx.sketch(10, 5, 0)


def _check_golden(r, case, printv):
    d = GOLDEN / case
    meta = load_meta(case)
    if meta["trace"]:
        np.testing.assert_allclose(r.trace, meta["trace"], atol=1.5e-6)
    if meta["em_trace"]:
        np.testing.assert_allclose(r.em_trace, meta["em_trace"], atol=1.5e-6)
    Uref = O.read_matrix(d / "eigvecs")
    np.testing.assert_allclose(O.sign_align(Uref, r.U), Uref, atol=2e-6)
    if printv:
        Vref = O.read_matrix(d / "loadings")
        np.testing.assert_allclose(O.sign_align(Vref, r.V), Vref, atol=2e-6)


def test_out_of_core_and_in_core_runs_share_one_context(ctx, golden_inputs):
    """The permute_plink order moves the allele frequencies with their SNPs and
    the in-core shuffle does not; runs of both kinds on one loaded data set
    must each restore the as-read state first."""
    say("out-of-core / in-core alternation on f4 (EMU, winSVD)")
    ctx.read_bed(golden_inputs["f4"])
    for case in ("f4_emu_d2_ooc", "f4_emu_d2", "f4_emu_d2_ooc", "f4_emu_d1_ooc"):
        params, printv = params_from_args(load_meta(case)["args"])
        say(f"  {case}")
        p = api.default_params(svd=params.svd, k=params.k, oversamples=params.oversamples, maxp=params.maxp,
                               tol=params.tol, bands=params.bands, emu=1, maxiter=params.maxiter,
                               tolem=params.tolem, memory=params.memory, precision="fp64")
        _check_golden(ctx.run(p, want_v=printv), case, printv)


def test_out_of_core_host_thin_q_without_flip(ctx, golden_inputs):
    """L = 60 takes the host thin-Q path; the out-of-core --svd 1 branch skips flipOmg."""
    say("out-of-core sSVD, L = 60, vs oracle")
    prefix = golden_inputs["f1"]
    packed, n, _ = O.read_plink(prefix)
    kw = dict(svd=1, k=30, oversamples=30, maxp=4, tol=0.0, memory=0.004)
    ref = O.run_pca_with_halko(O.Data.from_packed(packed, n, O.Params(**kw)))
    assert ref.epochs == 5
    ctx.read_bed(prefix)
    r = ctx.run(api.default_params(precision="fp64", **kw))
    np.testing.assert_allclose(r.trace, ref.trace, atol=1e-9)
    np.testing.assert_allclose(r.S, ref.S, rtol=1e-9)
    np.testing.assert_allclose(O.sign_align(ref.U, r.U), ref.U, atol=1e-7)


@pytest.mark.parametrize("memory", [0.0, 0.002])
def test_early_stop_matches_oracle(ctx, golden_inputs, memory):
    """--tol-rsvd 8e-4 stops f1 at epoch 13 of 21.  The next epoch's products
    are already enqueued when the stop rule fires; U, S, the loadings V (from
    the stopping epoch's G) and the trace must be the oracle's."""
    say(f"early stop, memory={memory}")
    prefix = golden_inputs["f1"]
    packed, n, _ = O.read_plink(prefix)
    kw = dict(svd=1, k=5, maxp=20, tol=8e-4, memory=memory)
    ref = O.run_pca_with_halko(O.Data.from_packed(packed, n, O.Params(**kw)))
    assert ref.epochs == 13
    ctx.read_bed(prefix)
    r = ctx.run(api.default_params(precision="fp64", **kw), want_v=True)
    assert r.info["epochs"] == 13
    np.testing.assert_allclose(r.trace, ref.trace, atol=1e-9)
    np.testing.assert_allclose(r.S, ref.S, rtol=1e-9)
    np.testing.assert_allclose(O.sign_align(ref.U, r.U), ref.U, atol=1e-7)
    Vref = np.empty_like(ref.V)
    Vref[ref.perm if ref.perm is not None else slice(None)] = ref.V
    np.testing.assert_allclose(O.sign_align(Vref, r.V), Vref, atol=1e-7)


def test_out_of_core_argument_errors(ctx, golden_inputs):
    say("out-of-core argument errors")
    ctx.read_bed(golden_inputs["f1"])
    with pytest.raises(api._lib.PcaoneError, match="only one block"):
        ctx.run(api.default_params(svd=1, k=5, maxp=1, memory=1.0, precision="fp64"))
    with pytest.raises(api._lib.PcaoneError, match="--maf"):
        ctx.run(api.default_params(svd=1, k=5, maxp=1, memory=0.002, maf=0.1, precision="fp64"))


@pytest.mark.parametrize("spec_kw", [dict(seed=9, nsamples=301, nsnps=777, npops=3),
                                     dict(seed=10, nsamples=5, nsnps=300, npops=2, freq_lo=0.01, freq_hi=0.05,
                                          fst_lo=0.05, fst_hi=0.15),
                                     dict(seed=11, nsamples=130, nsnps=400, npops=3, miss_rate=0.1,
                                          high_miss_frac=0.2, high_miss_rate=0.3)])
def test_device_generator_matches_host(ctx, spec_kw):
    from pcaone_amd import synth
    spec = synth.Spec(**spec_kw)
    say(f"device generator {spec_kw}")
    host, att = synth.packed(spec, with_attempts=True)
    ctx.generate(spec)
    dev = ctx.download_packed()
    assert np.array_equal(dev, host)
    if spec.nsamples == 5:
        assert att.max() > 0  # the redraw path was exercised


@pytest.mark.parametrize("form,wm,waves", [("bf16", 4, 8), ("bf16", 2, 8), ("bf16", 8, 4), ("fp32", 4, 8),
                                           ("bf16mix", 4, 8), ("bf16x1", 8, 4), ("bf16x1", 4, 8), ("fp16", 4, 8),
                                           ("fp16", 2, 8), ("fp16", 8, 4), ("int8", 2, 8), ("int8", 4, 8),
                                           ("int8", 4, 4), ("int8", 1, 4), ("fp8", 4, 8), ("fp8", 2, 8),
                                           ("fp8", 4, 4), ("fp8", 1, 4)])
def test_workgroup_shapes_agree(ctx, golden_inputs, form, wm, waves):
    """Every compiled workgroup shape gives the same decomposition (cfg1, 8 passes)."""
    say(f"shape form={form} wm={wm} waves={waves}")
    ctx.read_bed(golden_inputs["cfg1"])
    base = ctx.run(api.default_params(svd=1, k=10, maxp=7, tol=0.0, precision=form))
    r = ctx.run(api.default_params(svd=1, k=10, maxp=7, tol=0.0, precision=form, wm_g=wm, wm_h=wm,
                                   waves_g=waves, waves_h=waves))
    np.testing.assert_allclose(r.E, base.E, rtol=1e-9)
    assert O.pc_correlations(base.U, r.U).min() > 1 - 1e-9


# ---------------------------------------------------------------- output files (pcaone_run_bfile)
@pytest.mark.parametrize("case,form", [("f7_ld", "fp64"), ("f7_ld_std", "fp64"), ("f7_ld_d2", "fp64"),
                                       ("f7_ld_emu", "fp64"), ("f7_ld_ooc", "fp64"), ("f7_ld", "bf16"),
                                       ("f7_ld_emu", "bf16")])
def test_residual_files_match_reference(ctx, golden_inputs, case, form, tmp_path):
    """-D/--ld (src/Data.cpp:242-291): the run stays unscaled, the
    .residuals file (uint32 M, uint32 N, then N float32 per SNP in file order)
    holds A - U S V^T centered per SNP (the centered genotypes with
    --ld-stats 1), with the winSVD shuffle undone, the EMU imputation in place,
    and the out-of-core branch's unscaled, unimputed re-read; plus the .mbim.
    FP64 runs to 2e-6 (the reference's values are float32); BF16 to 1e-4."""
    meta = load_meta(case)
    params, printv = params_from_args(meta["args"])
    say(f"residual file {case} form={form} '{meta['args']}'")
    out = str(tmp_path / "run")
    p = api.default_params(svd=params.svd, k=params.k, oversamples=params.oversamples, maxp=params.maxp,
                           tol=params.tol, bands=params.bands, noshuffle=int(params.noshuffle), emu=int(params.emu),
                           maxiter=params.maxiter, tolem=params.tolem, memory=params.memory, ld=int(params.ld),
                           ld_stats=params.ld_stats, precision=form)
    ctx.run_bfile(golden_inputs[meta["input"]], out, p)
    d = GOLDEN / case
    R, Rref = O.read_residuals_file(out + ".residuals"), O.read_residuals_file(d / "residuals")
    assert R.shape == Rref.shape
    err = float(np.abs(R - Rref).max())
    say(f"  max |R - R_ref| {err:.2e}")
    assert err <= (2e-6 if form == "fp64" else 1e-4)
    mb, mb_ref = _rows(out + ".mbim"), _rows(d / "mbim")
    assert [r[:6] for r in mb] == [r[:6] for r in mb_ref]
    np.testing.assert_allclose(O.read_matrix(out + ".eigvals"), O.read_matrix(d / "eigvals"),
                               rtol=1e-5 if form == "fp64" else 1e-4)


def test_residuals_entry_point_matches_oracle(ctx, golden_inputs):
    """pcaone_residuals on the 2,000 x 20,000 input in row blocks (one not a
    multiple of the 32-row tile) against the oracle's write_residuals with the
    same U, S, V, after an unscaled FP64 run."""
    prefix = golden_inputs["cfg1"]
    params = O.Params(svd=1, k=10, maxp=4, tol=0.0, ld=True)
    packed, n, m = O.read_plink(prefix)
    data = O.Data.from_packed(packed, n, params)
    ref = O.run_pca_with_halko(data)
    ctx.read_bed(prefix)
    r = ctx.run(api.default_params(svd=1, k=10, maxp=4, tol=0.0, ld=1, precision="fp64"), want_v=True)
    blocks = [ctx.residuals(ref.U, ref.S, ref.V, j0, j1) for j0, j1 in ((0, 7001), (7001, 20000))]
    R = np.concatenate(blocks)
    err = float(np.abs(R - ref.residuals).max())
    say(f"residuals cfg1 k=10: max err {err:.2e}")
    assert err <= 2e-6
    np.testing.assert_allclose(r.S, ref.S, rtol=1e-9)
    R0 = ctx.residuals(None, None, None, 0, 100)  # --ld-stats 1: the centered matrix
    G = data.G[:, :100]
    np.testing.assert_allclose(R0, (G - G.mean(axis=0)).T.astype(np.float32), atol=1e-7)


def _rows(path):
    return [line.split("\t") for line in Path(path).read_text().splitlines()]


@pytest.mark.parametrize("case", ["f1_d1", "f1_d1_maf", "f1_d2", "f4_emu_d1", "f1_d2_ooc", "f4_emu_d2_ooc", "f1_d3",
                                  "f8_d0"])
def test_output_files_match_reference(ctx, golden_inputs, case, tmp_path):
    """pcaone_run_bfile writes the reference's files (src/Data.cpp:211-240,
    src/Utils.cpp:225-238): .sigvals with its '#N,M' header, .eigvals,
    .eigvecs, .eigvecs2 with the FID/IID columns of the .fam, and with -V the
    loadings and the .mbim (the .bim columns plus the allele frequency) in the
    original SNP order; numbers printed to 6 significant digits, tab-separated."""
    meta = load_meta(case)
    params, printv = params_from_args(meta["args"])
    say(f"output files {case} '{meta['args']}'")
    out = str(tmp_path / "run")
    p = api.default_params(svd=params.svd, k=params.k, oversamples=params.oversamples, maxp=params.maxp,
                           tol=params.tol, bands=params.bands, noshuffle=int(params.noshuffle), emu=int(params.emu),
                           maxiter=params.maxiter, tolem=params.tolem, maf=params.maf, memory=params.memory,
                           printv=int(printv), precision="fp64")
    ctx.run_bfile(golden_inputs[meta["input"]], out, p)
    d = GOLDEN / case
    sig, sig_ref = _rows(out + ".sigvals"), _rows(d / "sigvals")
    assert sig[0] == sig_ref[0]  # '#N,M'
    np.testing.assert_allclose(np.array(sig[1], float), np.array(sig_ref[1], float), rtol=1e-5)
    np.testing.assert_allclose(O.read_matrix(out + ".eigvals"), O.read_matrix(d / "eigvals"), rtol=1e-5)
    U, Uref = O.read_matrix(out + ".eigvecs"), O.read_matrix(d / "eigvecs")
    assert U.shape == Uref.shape
    np.testing.assert_allclose(O.sign_align(Uref, U), Uref, atol=2e-6)
    e2, e2_ref = _rows(out + ".eigvecs2"), _rows(d / "eigvecs2")
    assert e2[0] == e2_ref[0] and len(e2) == len(e2_ref)
    assert [r[:2] for r in e2] == [r[:2] for r in e2_ref]
    if printv:
        V, Vref = O.read_matrix(out + ".loadings"), O.read_matrix(d / "loadings")
        np.testing.assert_allclose(O.sign_align(Vref, V), Vref, atol=2e-6)
        mb, mb_ref = _rows(out + ".mbim"), _rows(d / "mbim")
        assert [r[:6] for r in mb] == [r[:6] for r in mb_ref]
        np.testing.assert_allclose([float(r[6]) for r in mb], [float(r[6]) for r in mb_ref], rtol=5e-6)
    for suffix in (".eigvals", ".eigvecs"):
        for line in Path(out + suffix).read_text().splitlines():
            for x in line.split("\t"):
                assert x == f"{float(x):.6g}", (suffix, x)
