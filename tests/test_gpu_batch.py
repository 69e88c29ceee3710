"""GPU parity of the batched form (many cohorts, one persistent launch per step,
the L x L steps on the device, one HIP graph per run).

Each cohort must reproduce what the single-matrix path (itself pinned against
the reference's golden outputs) returns for the same matrix and flags:
FP64 form to 1e-8 relative on the singular values and 1e-6 on the
sign-aligned PCs; BF16/FP32 forms to the accuracy bar against the FP64
result (eigenvalue relative error <= 1e-4, every PC correlated >= 0.9999).
"""
import numpy as np
import pytest

from conftest import GOLDEN, load_meta
from oracle import pcaone_oracle as O
from pcaone_amd import api, synth

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def ctx():
    c = api.Context(0)
    yield c
    c.close()


# uneven cohorts: N and M not multiples of the tiles, 2..4 populations
SPECS = [
    synth.Spec(seed=501, nsamples=300, nsnps=3000, npops=2),
    synth.Spec(seed=502, nsamples=1000, nsnps=20011, npops=3),
    synth.Spec(seed=503, nsamples=130, nsnps=1500, npops=2),
    synth.Spec(seed=504, nsamples=2047, nsnps=9000, npops=4),
    synth.Spec(seed=505, nsamples=777, nsnps=40000, npops=3),
]


def single_runs(ctx, specs, **kw):
    out = []
    for sp in specs:
        ctx.generate(sp)
        out.append(ctx.run(api.default_params(svd=1, **kw), want_v=True))
    return out


@pytest.fixture(scope="module")
def reference(ctx):
    say("batch reference: single-path FP64 runs")
    return single_runs(ctx, SPECS, k=5, oversamples=5, maxp=5, tol=0.0, precision="fp64")


def check_cohorts(ctx, res, ref, k, s_rtol, u_atol, v=False):
    for c, r in enumerate(ref):
        assert res["status"][c] == 0
        np.testing.assert_allclose(res["S"][c], r.S[:k], rtol=s_rtol)
        U = ctx.batch_get_u(c)
        np.testing.assert_allclose(O.sign_align(r.U, U), r.U, atol=u_atol)
        if v:
            V = ctx.batch_get_v(c)
            np.testing.assert_allclose(O.sign_align(r.V, V), r.V, atol=u_atol * 10)


def test_batch_fp64_matches_single_path(ctx, reference):
    say("batch fp64 eager vs single path")
    ctx.batch_generate(SPECS)
    res = ctx.batch_run(api.default_params(svd=1, k=5, oversamples=5, maxp=5, tol=0.0, precision="fp64"),
                        use_graph=False, want_v=True)
    assert list(res["epochs"]) == [6] * len(SPECS)
    check_cohorts(ctx, res, reference, 5, 1e-8, 1e-6, v=True)
    for c, sp in enumerate(SPECS):
        np.testing.assert_allclose(res["E"][c], res["S"][c] ** 2 / sp.nsnps)


@pytest.mark.parametrize("form", ["bf16", "fp32", "fp8"])
def test_batch_reduced_precision_meets_bar(ctx, reference, form):
    say(f"batch {form} vs fp64 single path")
    ctx.batch_generate(SPECS)
    res = ctx.batch_run(api.default_params(svd=1, k=5, oversamples=5, maxp=5, tol=0.0, precision=form))
    for c, r in enumerate(reference):
        rel = np.abs(res["E"][c] - r.E) / r.E
        corr = O.pc_correlations(r.U, ctx.batch_get_u(c))
        say(f"  cohort {c}: eig rel err {rel.max():.2e}, min corr {corr.min():.8f}")
        assert rel.max() <= 1e-4
        assert corr.min() >= 0.9999


def test_batch_fp8_matches_single_path_fp8(ctx):
    """The FP8 form batched: per-cohort power-of-two column scales of Omega
    (before its split) and of diag(sc) G~ (between the G epilogue and the
    split); every cohort equals the single path's FP8 run of its matrix up to
    the summation order of the FP32 partials."""
    say("batch fp8 vs single-path fp8")
    ref = single_runs(ctx, SPECS, k=5, oversamples=5, maxp=5, tol=0.0, precision="fp8")
    ctx.batch_generate(SPECS)
    res = ctx.batch_run(api.default_params(svd=1, k=5, oversamples=5, maxp=5, tol=0.0, precision="fp8"))
    for c, r in enumerate(ref):
        assert res["status"][c] == 0
        np.testing.assert_allclose(res["S"][c], r.S[:5], rtol=1e-6)
        assert O.pc_correlations(r.U, ctx.batch_get_u(c)).min() > 1 - 1e-6


@pytest.mark.parametrize("form", ["bf16", "fp8"])
def test_batch_graph_replays_match_eager(ctx, form):
    say(f"batch graph replays vs eager {form}")
    ctx.batch_generate(SPECS)
    p = api.default_params(svd=1, k=5, oversamples=5, maxp=4, tol=0.0, precision=form, wm_g=4, waves_g=8,
                           wm_h=4, waves_h=8, kchunk_g=2048, kchunk_h=4096)
    eager = ctx.batch_run(p, use_graph=False)
    U0 = [ctx.batch_get_u(c) for c in range(len(SPECS))]
    for rep in range(3):
        g = ctx.batch_run(p, use_graph=True)
        assert np.array_equal(g["S"], eager["S"]), f"replay {rep}"
        for c in range(len(SPECS)):
            assert np.array_equal(ctx.batch_get_u(c), U0[c]), f"replay {rep} cohort {c}"
        # eager work of the program between replays
        ctx.batch_run(p, use_graph=False)


def test_batch_stop_rule_matches_single_path(ctx):
    say("batch stop rule (tol 1e-4) vs single path")
    specs = SPECS[:3]
    ref = single_runs(ctx, specs, k=4, oversamples=4, maxp=20, tol=1e-4, precision="fp64")
    ctx.batch_generate(specs)
    res = ctx.batch_run(api.default_params(svd=1, k=4, oversamples=4, maxp=20, tol=1e-4, precision="fp64"))
    for c, r in enumerate(ref):
        assert res["epochs"][c] == r.info["epochs"]
        np.testing.assert_allclose(ctx.batch_trace(c), r.trace, atol=1e-9)
        np.testing.assert_allclose(res["S"][c], r.S, rtol=1e-8)


def test_batch_tuned_matches_fixed(ctx, reference):
    say("batch autotuned shapes vs fixed")
    ctx.batch_generate(SPECS)
    res = ctx.batch_run(api.default_params(svd=1, k=5, oversamples=5, maxp=5, tol=0.0, precision="fp64"))
    info = res["info"]
    say(f"  tuned: G {info['wm_g']}x{info['waves_g']} chunk {info['kchunk_g']}, "
        f"H {info['wm_h']}x{info['waves_h']} chunk {info['kchunk_h']}, {info['tuned_candidates']} candidates")
    assert info["tuned_candidates"] > 0
    check_cohorts(ctx, res, reference, 5, 1e-8, 1e-6)


@pytest.mark.parametrize("case", ["f1_d1", "cfg1_d1"])
def test_batch_matches_reference_golden(ctx, golden_inputs, case):
    """Cohorts read from the golden .bed files reproduce the reference binary's outputs."""
    meta = load_meta(case)
    say(f"batch golden {case} '{meta['args']}'")
    packed, n, m = O.read_plink(golden_inputs[meta["input"]])
    from conftest import params_from_args
    params, _ = params_from_args(meta["args"])
    # the same matrix twice plus a different cohort: each result is independent
    other = synth.packed(SPECS[0])
    ctx.batch_upload([(packed, n), (other, SPECS[0].nsamples), (packed, n)])
    res = ctx.batch_run(api.default_params(svd=1, k=params.k, oversamples=params.oversamples, maxp=params.maxp,
                                           tol=params.tol, precision="fp64"), use_graph=True)
    d = GOLDEN / case
    E = O.read_matrix(d / "eigvals").ravel()
    Uref = O.read_matrix(d / "eigvecs")
    for c in (0, 2):
        if meta["stop_epochs"]:
            assert res["epochs"][c] == meta["stop_epochs"][-1]
        np.testing.assert_allclose(res["E"][c], E, rtol=1e-5)
        np.testing.assert_allclose(O.sign_align(Uref, ctx.batch_get_u(c)), Uref, atol=2e-6)
    np.testing.assert_allclose(O.read_matrix(d / "sigvals").ravel(), res["S"][0], rtol=1e-5)


def test_batch_rejects_bad_input(ctx, golden_inputs):
    say("batch argument errors")
    packed, n, m = O.read_plink(golden_inputs["f4"])  # has missing genotypes
    with pytest.raises(api._lib.PcaoneError):
        ctx.batch_upload([(packed, n)])
    ctx.batch_generate(SPECS[:2])
    with pytest.raises(api._lib.PcaoneError):
        ctx.batch_run(api.default_params(svd=1, k=20, oversamples=20, precision="bf16"))  # L = 40 > 32
    with pytest.raises(api._lib.PcaoneError):
        ctx.batch_run(api.default_params(svd=2, k=5, precision="bf16"))
