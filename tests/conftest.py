import json
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
GOLDEN = ROOT / "tests" / "golden"
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X")


def _ensure_built():
    from oracle import build_oracle
    if not (ROOT / "oracle" / "_build" / "liboracle_rng.so").exists():
        build_oracle.build_rng()
    from pcaone_amd import build as pbuild
    if not pbuild.LIB.exists():
        pbuild.build()


@pytest.fixture(scope="session", autouse=True)
def built():
    _ensure_built()


def golden_cases():
    return sorted(p.name for p in GOLDEN.iterdir() if (p / "meta.json").exists())


def load_meta(case):
    return json.loads((GOLDEN / case / "meta.json").read_text())


@pytest.fixture(scope="session")
def golden_inputs(tmp_path_factory):
    """Prefix of every golden input; uncommitted ones are regenerated and hash-checked."""
    import hashlib
    from pcaone_amd import synth
    manifest = json.loads((GOLDEN / "inputs.json").read_text())
    tmp = tmp_path_factory.mktemp("golden_inputs")
    out = {}
    for name, m in manifest.items():
        if m["committed"]:
            prefix = GOLDEN / "data" / name
        else:
            prefix = tmp / name
            synth.write_plink(synth.Spec(**m["spec"]), str(prefix))
        digest = hashlib.sha256(Path(str(prefix) + ".bed").read_bytes()).hexdigest()
        assert digest == m["sha256"], f"golden input {name} does not regenerate bit-identically"
        out[name] = str(prefix)
    return out


@pytest.fixture(scope="session")
def golden_pgen(golden_inputs, tmp_path_factory):
    """case -> prefix of the .pgen/.pvar/.psam set the reference read for that
    golden case, written again from the committed .bed by pgen_writer and
    checked against the recorded SHA-256."""
    import hashlib
    sys.path.insert(0, str(GOLDEN))
    import pgen_writer
    tmp = tmp_path_factory.mktemp("golden_pgen")
    made = {}

    def get(case):
        if case not in made:
            meta = load_meta(case)
            opts = {k: v for k, v in meta["pgen"].items() if k != "sha256"}
            prefix = str(tmp / case)
            pgen_writer.bed_to_pgen(golden_inputs[meta["input"]], prefix, **opts)
            digest = hashlib.sha256(Path(prefix + ".pgen").read_bytes()).hexdigest()
            assert digest == meta["pgen"]["sha256"], f"{case}: .pgen does not regenerate bit-identically"
            made[case] = prefix
        return made[case]
    return get


def params_from_args(args: str, **extra):
    """Reference CLI flags (src/Cmd.cpp) -> oracle Params."""
    from oracle.pcaone_oracle import Params
    toks = args.split()
    kw = {}
    i = 0
    flags = {"-d": ("svd", int), "-k": ("k", int), "--maxp": ("maxp", int), "--rand": ("rand", int),
             "--seed": ("seed", int), "--tol-rsvd": ("tol", float), "--maf": ("maf", float), "-w": ("bands", int),
             "--maxiter": ("maxiter", int), "--oversamples": ("oversamples", int), "--tol-em": ("tolem", float),
             "--ld-stats": ("ld_stats", int),
             "-m": ("memory", float), "--buffer": ("buffer", int), "-n": ("threads", int)}
    printv = False
    while i < len(toks):
        t = toks[i]
        if t in flags:
            name, ty = flags[t]
            kw[name] = ty(toks[i + 1])
            i += 2
            continue
        if t == "-V":
            printv = True
        elif t == "-D":
            kw["ld"] = True
        elif t == "-S":
            kw["noshuffle"] = True
        elif t == "--emu":
            kw["emu"] = True
        elif t == "--hardcall":
            kw["hardcall"] = True
        else:
            raise ValueError(f"unknown flag {t}")
        i += 1
    kw.update(extra)
    return Params(**kw), printv
