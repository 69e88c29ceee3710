"""The reference-side binding (pcaone_amd/binding/pcaone_reference_binding.hpp)
compiled into the reference's own PLINK path (oracle/_ref/pcaone_ref_amd,
built by oracle/ref.mk from the reference's sources where they lie) and run on
the GPU: level 2 keeps the reference's computeUSV (host Householder QR,
fullPivHouseholderQr solve, JacobiSVD), EM loop and write_eigs_files around
the device computeGandH; level 1 is the whole run on the device.  The files it
writes must match the reference binary's golden files."""
import os
import subprocess
from pathlib import Path

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, load_meta
from oracle import pcaone_oracle as O

pytestmark = pytest.mark.gpu

BIN = ROOT / "oracle" / "_ref" / "pcaone_ref_amd"


@pytest.mark.skipif(not BIN.exists(), reason="oracle/_ref/pcaone_ref_amd is built only where the reference sources are")
@pytest.mark.parametrize("case,level", [("f1_d1", 2), ("f1_d2", 2), ("f4_emu_d1", 2), ("f6_d1", 2),
                                        ("f1_d1", 1), ("f1_d2_ooc", 1), ("f10_pg_d1", 2), ("f4_pg_emu", 2),
                                        ("f1_pg_ooc", 1)])
def test_binding_in_the_reference_path(golden_inputs, golden_pgen, case, level, tmp_path):
    meta = load_meta(case)
    out = str(tmp_path / "run")
    print(f"binding level {level}: {case} '{meta['args']}'", flush=True)
    env = dict(os.environ, PCAONE_AMD_LEVEL=str(level), PCAONE_AMD_PRECISION="0", OMP_NUM_THREADS="8")
    src = ["-p", golden_pgen(case)] if "pgen" in meta else ["-b", golden_inputs[meta["input"]]]
    cmd = [str(BIN), *src, *meta["args"].split(), "-o", out, "-v", "1"]
    res = subprocess.run(cmd, capture_output=True, text=True, env=env, timeout=120)
    assert res.returncode == 0, res.stdout[-2000:] + res.stderr[-2000:]
    d = GOLDEN / case
    np.testing.assert_allclose(O.read_matrix(out + ".eigvals"), O.read_matrix(d / "eigvals"), rtol=1e-5)
    np.testing.assert_allclose(O.read_matrix(out + ".sigvals"), O.read_matrix(d / "sigvals"), rtol=1e-5)
    U, Uref = O.read_matrix(out + ".eigvecs"), O.read_matrix(d / "eigvecs")
    np.testing.assert_allclose(O.sign_align(Uref, U), Uref, atol=2e-6)
    if (d / "loadings").exists() and Path(out + ".loadings").exists():
        V, Vref = O.read_matrix(out + ".loadings"), O.read_matrix(d / "loadings")
        np.testing.assert_allclose(O.sign_align(Vref, V), Vref, atol=2e-6)
    e2, e2_ref = Path(out + ".eigvecs2").read_text().splitlines(), (d / "eigvecs2").read_text().splitlines()
    assert e2[0] == e2_ref[0] and len(e2) == len(e2_ref)
    if meta["stop_epochs"] and level == 2:
        assert f"stops at epoch = {meta['stop_epochs'][-1]}" in " ".join(res.stdout.split()).replace("=  ", "= ")
