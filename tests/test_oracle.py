"""Pins the FP64 oracle (oracle/pcaone_oracle.py) before anything is checked against it."""
import numpy as np
import pytest

from conftest import GOLDEN, golden_cases, load_meta, params_from_args
from oracle import pcaone_oracle as O


# --------------------------------------------------------------- decode KATs
@pytest.mark.parametrize("n", [1021, 1022, 1023, 1024])
def test_decode_all_byte_values(n):
    row = np.arange(256, dtype=np.uint8)
    packed = row[None, :]
    codes = O.decode_codes(packed, n)
    assert codes.shape == (1, n)
    for j in range(n):
        assert codes[0, j] == (int(row[j // 4]) >> (2 * (j % 4))) & 3


def test_bed2geno_table():
    # 00 -> 2 copies of A1 (1.0), 01 -> missing, 10 -> het (0.5), 11 -> 0 copies
    codes = np.array([[0, 1, 2, 3]], dtype=np.uint8)
    assert list(O.BED2GENO[codes[0]]) == [1.0, -9.0, 0.5, 0.0]


def test_allele_freq_brute_force():
    rng = np.random.default_rng(0)
    codes = rng.integers(0, 4, size=(50, 37)).astype(np.uint8)
    codes[3] = 1  # all missing -> F = 0
    F = O.allele_freq(codes)
    for j in range(codes.shape[0]):
        vals = [O.BED2GENO[c] for c in codes[j] if c != 1]
        assert F[j] == (sum(vals) / len(vals) if vals else 0.0)


def test_centering_and_standardize_closed_form():
    codes = np.array([[0, 2, 3, 1, 2]], dtype=np.uint8)
    packed = np.zeros((1, 2), dtype=np.uint8)
    for j, c in enumerate(codes[0]):
        packed[0, j // 4] |= c << (2 * (j % 4))
    d = O.Data.from_packed(packed, 5, O.Params(svd=1, k=1))
    f = (1.0 + 0.5 + 0.0 + 0.5) / 4
    assert d.F[0] == f
    np.testing.assert_array_equal(d.G[:, 0], [1 - f, 0.5 - f, -f, 0.0, 0.5 - f])
    d.standardize_E()
    s = np.sqrt(2.0) / np.sqrt(f * (1 - f))
    np.testing.assert_allclose(d.G[:, 0], s * np.array([1 - f, 0.5 - f, -f, 0.0, 0.5 - f]), rtol=1e-15)


def test_standardize_skips_monomorphic():
    packed = np.full((1, 1), 0b11111111, dtype=np.uint8)
    d = O.Data.from_packed(packed, 4, O.Params(svd=1, k=1))
    d.standardize_E()
    assert d.F[0] == 0.0 and np.all(d.G == 0.0)


# --------------------------------------------------------------- small closed forms
def test_flip_omg_brute_force():
    rng = np.random.default_rng(1)
    a = rng.standard_normal((40, 6))
    b = a.copy()
    b[:, [1, 4]] *= -1
    b[:, 2] = rng.standard_normal(40)
    out = O.flip_omg(a, b)
    for i in range(6):
        flip = np.abs(a[:, i] - b[:, i]).sum() > 2 * np.abs(a[:, i] + b[:, i]).sum()
        np.testing.assert_array_equal(out[:, i], -b[:, i] if flip else b[:, i])
    np.testing.assert_array_equal(out[:, [1, 4]], a[:, [1, 4]])


def test_mev_closed_form():
    q = np.linalg.qr(np.random.default_rng(2).standard_normal((30, 4)))[0]
    assert abs(O.mev(q, q) - 1.0) < 1e-14
    assert abs(O.mev(q, -q) - 1.0) < 1e-14
    q2 = np.linalg.qr(np.random.default_rng(3).standard_normal((30, 8)))[0]
    assert abs(O.mev(q2[:, :4], q2[:, 4:])) < 1e-14


def test_flip_uv_makes_max_abs_loading_positive():
    rng = np.random.default_rng(4)
    U, V = rng.standard_normal((10, 3)), rng.standard_normal((20, 3))
    U2, V2 = O.flip_uv(U, V, False)
    for i in range(3):
        x = np.argmax(np.abs(V2[:, i]))
        assert V2[x, i] > 0
        np.testing.assert_array_equal(U2[:, i] * V2[:, i][0], U[:, i] * V[:, i][0])


def test_omega_is_deterministic_and_columnwise():
    a = O.omega(112, True, 5, 3)
    b = O.omega(112, True, 15, 1)
    np.testing.assert_array_equal(a.T.ravel(), b.ravel())  # column-major fill of one stream
    assert not np.array_equal(O.omega(113, True, 5, 3), a)
    u = O.omega(7, False, 100, 2)
    assert u.min() >= -1 and u.max() <= 1


def test_shuffle_is_permutation():
    p = O.shuffle_indices(1000)
    assert sorted(p.tolist()) == list(range(1000))
    np.testing.assert_array_equal(p, O.shuffle_indices(1000))


# --------------------------------------------------------------- golden (reference binary) runs
def _outputs_match(case, res, data, printv, tol_vec=2e-6):
    d = GOLDEN / case
    E = O.read_matrix(d / "eigvals").ravel() if (d / "eigvals").stat().st_size > 1 else np.zeros(0)
    if E.size:
        np.testing.assert_allclose(res.E, E, rtol=1e-5)
        S = O.read_matrix(d / "sigvals").ravel()
        np.testing.assert_allclose(res.S, S, rtol=1e-5)
    U = O.read_matrix(d / "eigvecs")
    assert U.shape == res.U.shape
    np.testing.assert_allclose(O.sign_align(U, res.U), U, atol=tol_vec)
    if data.params.svd == 3:  # flip_UV fixes the signs of the exact PCA (src/Main.cpp:215)
        np.testing.assert_allclose(res.U, U, atol=tol_vec)
    if printv and (d / "loadings").exists():
        V = O.read_matrix(d / "loadings")
        Vo = res.V
        if data.params.perm and data.perm is not None:
            Vo = np.empty_like(res.V)
            Vo[data.perm] = res.V
        np.testing.assert_allclose(O.sign_align(V, Vo), V, atol=tol_vec)
    if (d / "residuals").exists():
        R = O.read_residuals_file(d / "residuals")
        assert res.residuals is not None and res.residuals.shape == R.shape
        # float32 values; with --ld-stats 1 no PCs are involved and only the
        # summation order of the column means differs
        np.testing.assert_allclose(res.residuals, R, atol=1e-7 if data.params.ld_stats == 1 else 2e-6)


@pytest.mark.parametrize("case", golden_cases())
def test_oracle_matches_reference_binary(case, golden_inputs):
    """The oracle on the golden input (for a .pgen case: the .bed it was
    written from, read in the .pgen's -m order) against the reference binary."""
    meta = load_meta(case)
    params, printv = params_from_args(meta["args"], pgen="pgen" in meta)
    packed, n, _ = O.read_plink(golden_inputs[meta["input"]])
    data = O.Data.from_packed(packed, n, params)
    res = O.run_pca(data)
    if meta["trace"]:
        np.testing.assert_allclose(res.trace, meta["trace"], atol=1.5e-6)
    if meta["stop_epochs"]:
        assert res.epochs == meta["stop_epochs"][-1]
    if meta["em_trace"]:
        np.testing.assert_allclose(res.em_trace, meta["em_trace"], atol=1.5e-6)
    _outputs_match(case, res, data, printv)


def test_golden_mbim_allele_frequencies(golden_inputs):
    packed, n, m = O.read_plink(golden_inputs["f1"])
    F = O.allele_freq(O.decode_codes(packed, n))
    mb = [float(line.split("\t")[6]) for line in (GOLDEN / "f1_d1" / "mbim").read_text().splitlines()]
    np.testing.assert_allclose(F, mb, rtol=5e-6)


def _plant_no_flip(monkeypatch):
    monkeypatch.setattr(O, "flip_omg", lambda a, b: -b)


def _plant_omega_row_major(monkeypatch):
    orig = O.omega
    # the draw sequence laid out row by row instead of column by column
    monkeypatch.setattr(O, "omega", lambda s, g, n, l: orig(s, g, n, l).T.ravel().reshape(n, l).copy())


def _plant_no_scaling(monkeypatch):
    monkeypatch.setattr(O.Data, "standardize_E", lambda self: None)


def _plant_no_centering(monkeypatch):
    orig = O.Data.from_packed.__func__

    def bad(cls, packed, n, params):
        d = orig(cls, packed, n, params)
        d.G = np.where(d.G != 0.0, d.G + d.F[None, :], 0.0)
        return d
    monkeypatch.setattr(O.Data, "from_packed", classmethod(bad))


def _plant_swapped_clamp(monkeypatch):
    def bad(self, U, S, VT):
        pred = (U * S[None, :]) @ VT
        mask = self.C[self.perm].T if self.params.perm else self.C.T
        imp = np.minimum(np.maximum(pred, -(1.0 - self.F[None, :])), self.F[None, :])
        self.G = np.where(mask, imp, self.G)
    monkeypatch.setattr(O.Data, "fit_with_pi", bad)


def _residuals_with(monkeypatch, center=True, unpermute=True, subtract=True, standardized_ooc=False):
    def bad(data, S, U, VT):
        p = data.params
        G = (data.G0 if p.out_of_core and not standardized_ooc else data.G).copy()
        if subtract:
            G -= (U * S[None, :]) @ VT
        if center:
            G -= G.mean(axis=0, keepdims=True)
        out = np.empty((G.shape[1], G.shape[0]), dtype=np.float32)
        rows = data.perm if (unpermute and p.perm and data.perm is not None) else np.arange(G.shape[1])
        out[rows] = G.T.astype(np.float32)
        return out
    monkeypatch.setattr(O, "residuals", bad)


def _plant_residuals_uncentered(monkeypatch):
    _residuals_with(monkeypatch, center=False)


def _plant_residuals_not_unpermuted(monkeypatch):
    _residuals_with(monkeypatch, unpermute=False)


def _plant_residuals_not_subtracted(monkeypatch):
    _residuals_with(monkeypatch, subtract=False)


def _plant_residuals_ld_standardized(monkeypatch):
    # --ld runs the PCA on the standardized matrix (the flag ignored)
    orig = O.NormalOp.set_flags
    monkeypatch.setattr(O.NormalOp, "set_flags", lambda self, u, s: orig(self, u, True))


def _plant_full_no_flip(monkeypatch):
    monkeypatch.setattr(O, "flip_uv", lambda U, V, ubase=False: (U, V))


def _plant_full_gram_unscaled(monkeypatch):
    orig = O.run_pca_full

    def bad(data):
        r = orig(data)
        r.E = r.E * data.nsnps  # the Gram not divided by M
        return r
    monkeypatch.setattr(O, "run_pca_full", bad)


def _plant_iram_unstandardized(monkeypatch):
    monkeypatch.setattr(O.Data, "standardize_E", lambda self: None)


def _plant_ooc_band_factor_one(monkeypatch):
    orig = O.block_plan

    def bad(params, n, m):
        p = orig(params, n, m)
        p.band_factor = 1
        return p
    monkeypatch.setattr(O, "block_plan", bad)


def _plant_ooc_incore_shuffle(monkeypatch):
    monkeypatch.setattr(O, "permute_plink_order", lambda n, m, gb, w: O.shuffle_indices(m))


def _plant_pgen_one_thread(monkeypatch):
    orig = O.pgen_perm_order
    monkeypatch.setattr(O, "pgen_perm_order", lambda m, nb, bs, nt, seed: orig(m, nb, bs, 1, seed))


def _plant_pgen_default_seed(monkeypatch):
    # the partitions shuffled by a default-seeded engine, as the in-core shuffle draws
    orig = O.pgen_perm_order
    monkeypatch.setattr(O, "pgen_perm_order", lambda m, nb, bs, nt, seed: orig(m, nb, bs, nt, 1))


def _plant_pgen_plink_order(monkeypatch):
    monkeypatch.setattr(O, "pgen_perm_order",
                        lambda m, nb, bs, nt, seed: O.permute_plink_order(200, m, 2, nb))


def _plant_ooc_no_flip(monkeypatch):
    monkeypatch.setattr(O, "flip_omg", lambda a, b: b)


def _plant_ooc_impute_unclamped(monkeypatch):
    orig = O.Data.block_matrix

    def bad(self, update, standardize, U, S, VT):
        A = orig(self, False, False, U, S, VT)
        if update and self.params.emu:
            A = np.where(self.C.T, (U * S[None, :]) @ VT, A)
        if standardize:
            sd = np.sqrt(self.F * (1.0 - self.F))
            A = A * np.where(sd > O.VAR_TOL, np.sqrt(2.0) / np.where(sd > O.VAR_TOL, sd, 1.0), 1.0)[None, :]
        return A
    monkeypatch.setattr(O.Data, "block_matrix", bad)


@pytest.mark.parametrize("case", [c for c in golden_cases() if "block_plan" in load_meta(c)])
def test_block_plan_matches_reference_log(case):
    """Data::prepare's plan (src/Data.cpp:42-84) as the reference binary logged it."""
    meta = load_meta(case)
    params, _ = params_from_args(meta["args"])
    n, m = {"f1": (200, 2000), "f4": (300, 3000), "f7": (64, 500)}[meta["input"]]
    plan = O.block_plan(params, n, m)
    assert [plan.blocksize, plan.nblocks, plan.band_factor] == meta["block_plan"]
    assert plan.start[-1] < m <= plan.stop[-1] + 1


@pytest.mark.parametrize("case,plant", [
    ("f1_d2_ooc", _plant_ooc_band_factor_one),
    ("f1_d2_ooc", _plant_ooc_incore_shuffle),
    ("f1_d2_ooc", _plant_ooc_no_flip),
    ("f1_pg_ooc", _plant_pgen_one_thread),
    ("f1_pg_ooc", _plant_pgen_default_seed),
    ("f1_pg_ooc", _plant_pgen_plink_order),
    ("f4_emu_d1_ooc", _plant_ooc_impute_unclamped),
    ("f1_d2", _plant_no_flip),
    ("f1_d1_p0", _plant_omega_row_major),
    ("f1_d1", _plant_no_scaling),
    ("f1_d1", _plant_no_centering),
    ("f4_emu_d1", _plant_swapped_clamp),
    ("f7_ld_emu", _plant_residuals_uncentered),  # imputed entries move the column means
    ("f7_ld_d2", _plant_residuals_not_unpermuted),
    ("f7_ld", _plant_residuals_not_subtracted),
    ("f7_ld", _plant_residuals_ld_standardized),
    ("f1_d3", _plant_full_no_flip),
    ("f1_d3", _plant_full_gram_unscaled),
    ("f1_d0", _plant_iram_unstandardized),
])
def test_planted_mistake_is_caught(case, plant, golden_inputs, monkeypatch):
    """A plausible mistake planted in the oracle must fail the golden comparison."""
    meta = load_meta(case)
    params, printv = params_from_args(meta["args"], pgen="pgen" in meta)
    packed, n, _ = O.read_plink(golden_inputs[meta["input"]])
    plant(monkeypatch)
    data = O.Data.from_packed(packed, n, params)
    res = O.run_pca(data)
    with pytest.raises(AssertionError):
        if meta["trace"]:
            np.testing.assert_allclose(res.trace, meta["trace"], atol=1.5e-6)
        if meta["em_trace"]:
            np.testing.assert_allclose(res.em_trace, meta["em_trace"], atol=1.5e-6)
        _outputs_match(case, res, data, printv)


def test_dense_oracle_agrees_at_convergence(golden_inputs):
    """cfg1 (2,000 x 20,000, 3 populations): the two structured PCs of the RSVD
    match the exact SVD of the standardized matrix."""
    meta = load_meta("cfg1_d1")
    params, _ = params_from_args(meta["args"])
    packed, n, _ = O.read_plink(golden_inputs["cfg1"])
    data = O.Data.from_packed(packed, n, params)
    data.standardize_E()
    U, S, E, V = O.dense_pca(data, params.k)
    Eref = O.read_matrix(GOLDEN / "cfg1_d1" / "eigvals").ravel()
    Uref = O.read_matrix(GOLDEN / "cfg1_d1" / "eigvecs")
    np.testing.assert_allclose(Eref[:2], E[:2], rtol=1e-5)
    assert np.all(O.pc_correlations(U[:, :2], Uref[:, :2]) > 0.9999)


# ---------------------------------------------------------------- metric helpers (closed forms)
def test_full_piv_solve_closed_forms():
    """The restated fullPivHouseholderQr solve: the ordinary solve for an
    invertible matrix; for a singular consistent system an exact solution with
    exactly n - rank unknowns set to zero (the basic solution); a tiny planted
    pivot below eps * n of the largest counts as zero."""
    rng = np.random.default_rng(7)
    A = rng.standard_normal((12, 12))
    B = rng.standard_normal((12, 5))
    np.testing.assert_allclose(O.full_piv_solve(A, B), np.linalg.solve(A, B), atol=1e-12)
    A2 = rng.standard_normal((12, 7)) @ rng.standard_normal((7, 12))
    B2 = A2 @ rng.standard_normal((12, 3))
    X = O.full_piv_solve(A2, B2)
    np.testing.assert_allclose(A2 @ X, B2, atol=1e-11)
    assert int((np.abs(X).sum(axis=1) == 0).sum()) == 12 - 7
    # upper triangular with one pivot at 1e-18 of the others: that unknown is 0
    R = np.triu(rng.standard_normal((6, 6))) + 4 * np.eye(6)
    R[5, 5] = 1e-18
    X3 = O.full_piv_solve(R.T, rng.standard_normal((6, 2)))
    assert np.all(X3[5] == 0.0) and np.all(np.abs(X3[:5]) > 0)
    # the same with LAPACK's solve blows that unknown up instead
    assert np.abs(np.linalg.solve(R.T, rng.standard_normal((6, 2)))[5]).max() > 1e10


def test_pc_correlations_closed_forms():
    """|Pearson r| per column: invariant to sign and to a positive scale and
    shift (1), zero for a column orthogonal to the other after centering,
    and cos(theta) for two centered unit vectors at angle theta."""
    rng = np.random.default_rng(8)
    x = rng.standard_normal((500, 3))
    assert np.allclose(O.pc_correlations(x, x), 1.0, atol=1e-14)
    assert np.allclose(O.pc_correlations(x, -3.0 * x + 7.0), 1.0, atol=1e-14)
    a = x[:, :1] - x[:, :1].mean()
    b = rng.standard_normal((500, 1))
    b -= b.mean()
    b -= a * (a.T @ b).item() / (a.T @ a).item()  # centered, orthogonal to a
    assert O.pc_correlations(a, b)[0] < 1e-14
    u = a / np.linalg.norm(a)
    v = b / np.linalg.norm(b)
    for theta in (0.1, 0.7, 1.3, 2.5):
        w = np.cos(theta) * u + np.sin(theta) * v
        assert abs(O.pc_correlations(u, w)[0] - abs(np.cos(theta))) < 1e-13
    # column pairs are matched by index, not by best match
    y = x[:, ::-1].copy()
    r = O.pc_correlations(x, y)
    assert abs(r[1] - 1.0) < 1e-14 and r[0] < 0.2 and r[2] < 0.2


def test_principal_angles_closed_forms():
    """Angles between the column spans: zero for the same span under any
    invertible mixing, pi/2 for orthogonal spans, and the planted angles of
    two planes built from orthonormal directions."""
    rng = np.random.default_rng(9)
    Q = np.linalg.qr(rng.standard_normal((300, 6)))[0]
    A = Q[:, :3]
    mix = rng.standard_normal((3, 3)) + 3 * np.eye(3)
    assert np.max(O.principal_angles(A, A @ mix)) < 1e-7
    assert np.allclose(O.principal_angles(A, Q[:, 3:]), np.pi / 2, atol=1e-12)
    t = np.array([0.05, 0.4, 1.2])
    B = A * np.cos(t) + Q[:, 3:] * np.sin(t)  # column i turned by t_i towards its own orthogonal direction
    got = np.sort(O.principal_angles(A, B))
    assert np.allclose(got, np.sort(t), atol=1e-12)
