"""The --pgen path on the device against the reference binary's golden runs:
the output files of pcaone_run_pgen (.mbim from the .pvar, .eigvecs2 from the
.psam), streamed -m runs that decode the .pgen at every block load, and the
allele frequencies of a .pgen equal to those of the .bed it holds
(src/FilePgen.cpp:27-287, src/Main.cpp:150-231, src/Utils.cpp:240-265,405-430)."""
from pathlib import Path

import numpy as np
import pytest

from conftest import GOLDEN, golden_cases, load_meta, params_from_args
from oracle import pcaone_oracle as O
from pcaone_amd import api
from pcaone_amd._lib import PcaoneError

pytestmark = pytest.mark.gpu

PGEN_CASES = [c for c in golden_cases() if "pgen" in load_meta(c)]


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def ctx():
    c = api.Context(0)
    yield c
    c.close()


def _rows(path):
    return [line.split("\t") for line in Path(path).read_text().splitlines()]


def _params(meta, **kw):
    params, printv = params_from_args(meta["args"], pgen=True)
    p = api.default_params(svd=params.svd, k=params.k, oversamples=params.oversamples, maxp=params.maxp,
                           tol=params.tol, bands=params.bands, noshuffle=int(params.noshuffle), emu=int(params.emu),
                           maxiter=params.maxiter, tolem=params.tolem, maf=params.maf, memory=params.memory,
                           threads=params.threads, hardcall=int(params.hardcall), printv=int(printv), **kw)
    return p, printv


@pytest.mark.parametrize("case", PGEN_CASES)
def test_pgen_output_files_match_reference(ctx, golden_pgen, case, tmp_path):
    """pcaone_run_pgen writes the reference's files: '#N,M' .sigvals, .eigvals,
    .eigvecs, the .eigvecs2 FID/IID columns from the .psam (FID 0 when it has
    no FID column), and with -V the loadings and the .mbim (each .pvar line as
    a .bim line plus the allele frequency) in the file's SNP order."""
    meta = load_meta(case)
    p, printv = _params(meta, precision="fp64")
    say(f"run_pgen {case} '{meta['args']}' {meta['pgen']}")
    out = str(tmp_path / "run")
    ctx.run_pgen(golden_pgen(case), out, p)
    d = GOLDEN / case
    sig, sig_ref = _rows(out + ".sigvals"), _rows(d / "sigvals")
    assert sig[0] == sig_ref[0]
    np.testing.assert_allclose(O.read_matrix(out + ".eigvals"), O.read_matrix(d / "eigvals"), rtol=1e-5)
    U, Uref = O.read_matrix(out + ".eigvecs"), O.read_matrix(d / "eigvecs")
    np.testing.assert_allclose(O.sign_align(Uref, U), Uref, atol=2e-6)
    e2, e2_ref = _rows(out + ".eigvecs2"), _rows(d / "eigvecs2")
    assert [r[:2] for r in e2] == [r[:2] for r in e2_ref]
    if printv:
        V, Vref = O.read_matrix(out + ".loadings"), O.read_matrix(d / "loadings")
        np.testing.assert_allclose(O.sign_align(Vref, V), Vref, atol=2e-6)
        mb, mb_ref = _rows(out + ".mbim"), _rows(d / "mbim")
        assert [r[:6] for r in mb] == [r[:6] for r in mb_ref]
        np.testing.assert_allclose([float(r[6]) for r in mb], [float(r[6]) for r in mb_ref], rtol=5e-6)


@pytest.mark.parametrize("case", ["f1_pg_ooc", "f4_pg_emu_ooc"])
def test_pgen_streamed_runs_match_reference(ctx, golden_pgen, case, tmp_path):
    """The -m runs with the .pgen streamed: every block load decodes the
    block's records (in the .pgen shuffle's order) instead of reading the
    resident matrix; results equal the reference's."""
    meta = load_meta(case)
    p, _ = _params(meta, precision="fp64", stream=1)
    say(f"run_pgen streamed {case} '{meta['args']}'")
    out = str(tmp_path / "run")
    info = ctx.run_pgen(golden_pgen(case), out, p)
    d = GOLDEN / case
    np.testing.assert_allclose(O.read_matrix(out + ".eigvals"), O.read_matrix(d / "eigvals"), rtol=1e-5)
    U, Uref = O.read_matrix(out + ".eigvecs"), O.read_matrix(d / "eigvecs")
    np.testing.assert_allclose(O.sign_align(Uref, U), Uref, atol=2e-6)
    if meta["em_trace"]:
        assert info["em_iters"] == len(meta["em_trace"])


def test_pgen_allele_freq_equals_bed(ctx, golden_inputs, golden_pgen):
    """read_pgen of the LD input (every record type) gives the allele
    frequencies read_bed gives for the .bed it was written from, bit for bit."""
    ctx.read_bed(golden_inputs["f10"])
    Fb = ctx.allele_freq()
    n, m = ctx.read_pgen(golden_pgen("f10_pg_d1"))
    assert (n, m) == (700, 1200)
    np.testing.assert_array_equal(ctx.allele_freq(), Fb)
    np.testing.assert_array_equal(ctx.download_packed(), O.read_plink(golden_inputs["f10"])[0])


def test_pgen_dosage_file_needs_hardcall(ctx, golden_pgen, tmp_path):
    """A .pgen that stores dosages is read only with hardcall (the reference
    reads dosages without --hardcall; this reader takes hardcalls only)."""
    prefix = golden_pgen("f1_pg_dos")
    p, _ = _params(load_meta("f1_pg_dos"), precision="fp64")
    p.hardcall = 0
    with pytest.raises(PcaoneError) as e:
        ctx.run_pgen(prefix, str(tmp_path / "x"), p)
    assert e.value.status == 4 and "hardcall" in str(e.value)
