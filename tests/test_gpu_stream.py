"""Streamed genotypes (data beyond HBM): the packed .bed rows stay on the host
and every product loads its block of the -m plan into block-sized device
layouts, as the reference's out-of-core runs read their blocks on every pass
(src/FilePlink.cpp:122-298, src/Halko.cpp:129-152,225-268).  The six golden
-m cases of the reference binary run through this path, from the mapped
.bed file and from host rows, in FP64 (to the reference's printed precision)
and BF16 (to the reduced-precision bar)."""
import numpy as np
import pytest

from conftest import GOLDEN, load_meta, params_from_args
from oracle import pcaone_oracle as O
from pcaone_amd import api, synth
from pcaone_amd._lib import PcaoneError

pytestmark = pytest.mark.gpu

OOC = ["f1_d1_ooc", "f1_d1_ooc_min", "f1_d2_ooc", "f1_d2_ooc_noshuf", "f4_emu_d1_ooc", "f4_emu_d2_ooc"]


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def ctx():
    c = api.Context(0)
    yield c
    c.close()


def _params(params, form, **extra):
    return api.default_params(svd=params.svd, k=params.k, oversamples=params.oversamples, maxp=params.maxp,
                              tol=params.tol, bands=params.bands, noshuffle=int(params.noshuffle),
                              emu=int(params.emu), maxiter=params.maxiter, tolem=params.tolem,
                              memory=params.memory, buffer=params.buffer, precision=form, **extra)


@pytest.mark.parametrize("case", OOC)
def test_streamed_bed_matches_reference(ctx, golden_inputs, case, tmp_path):
    meta = load_meta(case)
    params, printv = params_from_args(meta["args"])
    say(f"streamed {case} '{meta['args']}'")
    out = str(tmp_path / "run")
    info = ctx.run_bfile(golden_inputs[meta["input"]], out, _params(params, "fp64", stream=1, printv=int(printv)))
    d = GOLDEN / case
    if meta["trace"]:
        np.testing.assert_allclose(ctx._trace(0, info["n_trace"]), meta["trace"], atol=1.5e-6)
    if meta["em_trace"]:
        np.testing.assert_allclose(ctx._trace(1, max(info["em_iters"], 0)), meta["em_trace"], atol=1.5e-6)
    np.testing.assert_allclose(O.read_matrix(out + ".eigvals"), O.read_matrix(d / "eigvals"), rtol=1e-5)
    U, Uref = O.read_matrix(out + ".eigvecs"), O.read_matrix(d / "eigvecs")
    np.testing.assert_allclose(O.sign_align(Uref, U), Uref, atol=2e-6)
    if printv and (d / "loadings").exists():
        V, Vref = O.read_matrix(out + ".loadings"), O.read_matrix(d / "loadings")
        np.testing.assert_allclose(O.sign_align(Vref, V), Vref, atol=2e-6)
    # every pass loaded every block of the plan again
    nb = meta["block_plan"][1]
    say(f"  blocks per pass {nb}")


@pytest.mark.parametrize("case,form", [("f1_d2_ooc", "bf16"), ("f4_emu_d2_ooc", "bf16"), ("f1_d1_ooc", "fp32")])
def test_streamed_host_rows_reduced_precision(ctx, golden_inputs, case, form):
    meta = load_meta(case)
    params, printv = params_from_args(meta["args"])
    say(f"streamed host rows {case} {form}")
    packed, n, m = O.read_plink(golden_inputs[meta["input"]])
    plan = O.block_plan(params, n, m)
    ctx.stream_host(packed, n, plan.blocksize)
    r = ctx.run(_params(params, form), want_v=printv)
    ref = O.run_plink(golden_inputs[meta["input"]], params)
    rel = np.abs(r.E - ref.E) / ref.E
    corr = O.pc_correlations(ref.U, r.U)
    say(f"  eig rel err max {rel.max():.2e}  min PC corr {corr.min():.8f}")
    assert rel.max() <= 1e-4 and corr.min() >= 0.9999


@pytest.mark.parametrize("svd,bands,memory,nblocks", [(1, 64, 3, 3), (2, 4, 2, 4)])
def test_streamed_blocks_of_several_chunks_match_resident(ctx, svd, bands, memory, nblocks):
    """A block load stages its rows through two pinned slots of about 32 MB,
    so blocks of 40,000 samples (10,000 bytes a row, 3,355 rows a slot) load
    in two to four chunks, the last one ragged; the winSVD case gathers its
    rows in the shuffled order across the chunk bounds.  Streamed runs equal
    the resident run of the same rows in FP64."""
    spec = synth.Spec(seed=23, nsamples=40000, nsnps=24000, miss_rate=0.02)
    packed = synth.packed(spec)
    p = api.default_params(svd=svd, k=5, maxp=3, bands=bands, memory=memory, precision="fp64")
    plan = O.block_plan(p, spec.nsamples, spec.nsnps)
    say(f"several chunks: svd {svd} bands {bands} -m {memory}: {plan.nblocks} blocks of {plan.blocksize}")
    assert plan.nblocks == nblocks and plan.blocksize * spec.bytes_per_snp > 32 << 20
    ctx.stream_host(packed, spec.nsamples, plan.blocksize)
    rs = ctx.run(p)
    assert rs.info["stream_loads"] >= 2 * nblocks  # every pass loads every block
    ctx.upload_packed(packed, spec.nsamples)
    rr = ctx.run(p)
    assert rr.info["stream_loads"] == 0
    rel = np.abs(rs.E - rr.E) / rr.E
    corr = O.pc_correlations(rr.U, rs.U)
    say(f"  eig rel diff max {rel.max():.2e}  min PC corr {corr.min():.12f}")
    assert rel.max() <= 1e-9 and corr.min() >= 1 - 1e-9


def test_streamed_argument_errors(ctx, golden_inputs):
    packed, n, m = O.read_plink(golden_inputs["f1"])
    ctx.stream_host(packed, n, 500)
    # an in-core run needs the matrix resident
    with pytest.raises(PcaoneError, match="out-of-core"):
        ctx.run(api.default_params(svd=1, k=5, maxp=2, precision="fp64"))
    # plan blocks larger than the streamed block
    with pytest.raises(PcaoneError, match="exceed the streamed block"):
        ctx.run(api.default_params(svd=1, k=5, maxp=2, memory=0.002, precision="fp64"))
    # resident-only operations
    with pytest.raises(PcaoneError, match="resident"):
        ctx.run(api.default_params(svd=3, k=5, precision="fp64"))
