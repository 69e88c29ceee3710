"""The .pgen reader without a GPU: every storage mode and record type decoded
bit-exactly into the .bed rows the file was written from, the .pgen -m shuffle
against the oracle, and the reader's error paths (src/FilePgen.cpp,
src/FilePgen.hpp, external/pgenlib/include/pgenlib_read.cc)."""
import itertools
import struct
import sys
from pathlib import Path

import numpy as np
import pytest

from conftest import GOLDEN, load_meta
from oracle import pcaone_oracle as O
from pcaone_amd import api, synth
from pcaone_amd._lib import PcaoneError

sys.path.insert(0, str(GOLDEN))
import pgen_writer as W  # noqa: E402

ERR_IO = 4


def _bed_rows(prefix):
    raw = Path(str(prefix) + ".bed").read_bytes()
    n = len(Path(str(prefix) + ".fam").read_text().splitlines())
    return np.frombuffer(raw[3:], dtype=np.uint8).reshape(-1, (n + 3) // 4), n


def _synth_set(tmp_path, name, **kw):
    spec = synth.Spec(**kw)
    prefix = str(tmp_path / name)
    synth.write_plink(spec, prefix)
    return prefix


WRITES = [
    dict(mode="plink1"),
    dict(mode="fixed"),
    dict(mode="fixed", nonref=3),
    dict(mode="var", policy="raw"),
    dict(mode="var", policy="cycle"),
    dict(mode="var", policy="plink2", nonref=3),
    dict(mode="var", policy="cycle", len_bytes=4, nonref=2),
    dict(mode="var", policy="cycle", dosage=True),
]


@pytest.mark.parametrize("opts", WRITES, ids=lambda o: "-".join(f"{k}{v}" for k, v in o.items()))
@pytest.mark.parametrize("inp", ["f10", "f5", "f4"])
def test_decode_matches_bed(inp, opts, golden_inputs, tmp_path):
    """pcaone_pgen_decode of a .pgen written from a golden .bed gives back the
    .bed rows byte for byte (trailing bits zero), whole and in a middle range."""
    src = golden_inputs[inp]
    rows, n = _bed_rows(src)
    prefix = str(tmp_path / "x")
    W.bed_to_pgen(src, prefix, **opts)
    hard = bool(opts.get("dosage"))
    assert api.pgen_dims(prefix) == (n, rows.shape[0])
    np.testing.assert_array_equal(api.pgen_decode(prefix, hardcall=hard), rows)
    np.testing.assert_array_equal(api.pgen_decode(prefix, 101, 377, hardcall=hard), rows[101:377])


def test_writer_uses_every_record_type(golden_inputs):
    """The 'cycle' files of the LD input hold every main-track record type, so
    the decode tests above cover each one."""
    rows, n = _bed_rows(golden_inputs["f10"])
    g = W.bed_to_geno(rows, n)
    seen, ld = set(), None
    for j in range(g.shape[0]):
        vt, _ = W.encode_variant(g[j], ld, j, "cycle")
        if vt not in (2, 3):
            ld = g[j]
        seen.add(vt)
    assert seen == set(range(8))


@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 63, 255, 256, 257, 1029])
def test_decode_ragged_sample_counts(n, tmp_path):
    """Sample counts around the byte, 8-bit-index and difference-list group
    boundaries (1-byte sample indices up to 255, 2 bytes from 256)."""
    rng = np.random.default_rng(40 + n)
    # genotypes drawn directly: mostly 0 copies, so the sparse record types fit
    alt = rng.choice(4, size=(300, n), p=[0.85, 0.08, 0.04, 0.03]).astype(np.uint8)
    inv = np.array([3, 2, 0, 1], dtype=np.uint8)
    rows = np.stack([np.frombuffer(W.pack2(inv[r]), dtype=np.uint8) for r in alt])
    prefix = str(tmp_path / "s")
    Path(prefix + ".bed").write_bytes(b"l\x1b\x01" + rows.tobytes())
    Path(prefix + ".bim").write_text("".join(f"1\ts{j}\t0\t{j + 1}\tA\tC\n" for j in range(300)))
    Path(prefix + ".fam").write_text("".join(f"f{i} i{i} 0 0 0 -9\n" for i in range(n)))
    for opts in WRITES[1:6]:
        W.bed_to_pgen(prefix, prefix + "_p", **opts)
        np.testing.assert_array_equal(api.pgen_decode(prefix + "_p"), rows)


def test_decode_across_variant_blocks(tmp_path):
    """70,000 variants: two blocks of the variable-width index (65,536
    variants each), records of every type on both sides of the boundary."""
    prefix = _synth_set(tmp_path, "big", seed=77, nsamples=24, nsnps=70000, npops=2, miss_rate=0.05)
    rows, _ = _bed_rows(prefix)
    W.bed_to_pgen(prefix, prefix + "_p", mode="var", policy="cycle")
    np.testing.assert_array_equal(api.pgen_decode(prefix + "_p"), rows)
    np.testing.assert_array_equal(api.pgen_decode(prefix + "_p", 65530, 65540), rows[65530:65540])


def test_psam_counts_data_lines_only(golden_inputs, tmp_path):
    """N is the number of .psam lines that do not start with '#'
    (src/FilePgen.hpp:18-25), with or without a FID column."""
    src = golden_inputs["f1"]
    for fid in (True, False):
        W.bed_to_pgen(src, str(tmp_path / "p"), mode="fixed", fid=fid)
        assert api.pgen_dims(str(tmp_path / "p")) == (200, 2000)


def _expect_io(fn, *args, match=None):
    with pytest.raises(PcaoneError) as e:
        fn(*args)
    assert e.value.status == ERR_IO
    if match:
        assert match in str(e.value)


def test_reader_errors(golden_inputs, tmp_path):
    """Bad magic, sample-major .bed payload, a .psam that disagrees with the
    header, stored allele counts (rejected by the reference's loader,
    external/pgenlib/pgenlibr.cpp:65-67), dosages without hardcall, truncated
    records, an unsupported storage mode."""
    src = golden_inputs["f1"]
    p = str(tmp_path / "e")
    W.bed_to_pgen(src, p, mode="var", policy="cycle")
    good = Path(p + ".pgen").read_bytes()

    def write(b):
        Path(p + ".pgen").write_bytes(b)

    write(b"\x00\x1b" + good[2:])
    _expect_io(api.pgen_dims, p, match="magic")
    write(b"l\x1b\x00" + good[3:])
    _expect_io(api.pgen_dims, p, match="sample-major")
    write(good[:7] + struct.pack("<I", 199) + good[11:])
    _expect_io(api.pgen_dims, p, match="samples")
    write(good[:11] + bytes([good[11] | 0x10]) + good[12:])
    _expect_io(api.pgen_dims, p, match="allele count")
    write(good[:2] + b"\x05" + good[3:])
    _expect_io(api.pgen_dims, p, match="storage mode")
    write(good[:-40])
    _expect_io(api.pgen_dims, p, match="past the end")
    W.bed_to_pgen(src, p, mode="var", policy="cycle", dosage=True)
    _expect_io(api.pgen_decode, p, 0, 10, False, match="hardcall")
    assert api.pgen_decode(p, 0, 10, hardcall=True).shape == (10, 50)
    W.bed_to_pgen(src, p, mode="fixed")
    fixed = Path(p + ".pgen").read_bytes()
    write(fixed[:-1])
    _expect_io(api.pgen_dims, p, match="file size")


def test_malformed_difference_list_is_an_error(golden_inputs, tmp_path):
    """A difference-list sample index past N fails the decode (no write out of
    bounds)."""
    src = golden_inputs["f5"]
    rows, n = _bed_rows(src)
    g = W.bed_to_geno(rows, n)[:4].copy()
    g[0] = 0
    g[0, [3, 9]] = 1  # record type 4 with two entries
    p = str(tmp_path / "m")
    W.write_pgen(p + ".pgen", g, mode="var", policy="plink2")
    Path(p + ".psam").write_text("#IID\n" + "".join(f"i{i}\n" for i in range(n)))
    raw = bytearray(Path(p + ".pgen").read_bytes())
    # the first record starts at the offset stored after the 12-byte header
    at = struct.unpack_from("<Q", raw, 12)[0]
    assert raw[at] == 2 and raw[at + 1] == 3  # length 2, first index 3
    raw[at + 1] = 200  # first index beyond N = 150
    Path(p + ".pgen").write_bytes(bytes(raw))
    _expect_io(api.pgen_decode, p, 0, 1, False, match="malformed")


def test_pgen_perm_order_matches_oracle():
    """compute_pgen_perm (src/FilePgen.cpp:91-151): the library's order equals
    the oracle's over a grid of sizes, batches, block sizes, threads and seeds,
    and is a permutation."""
    for m, nb, bs, nt, seed in itertools.product([1, 9, 100, 2001], [1, 4, 16], [1, 3, 42, 500], [1, 4, 13],
                                                 [112, 7]):
        a = api.pgen_perm_order(m, nb, bs, nt, seed)
        np.testing.assert_array_equal(a, O.pgen_perm_order(m, nb, bs, nt, seed))
        assert sorted(a.tolist()) == list(range(m))


def test_pgen_perm_order_hand_worked():
    """One batch, blocks of 4 over 2 threads: thread 0 owns sources 0..3 and
    thread 1 sources 4..7 (before the shuffle); block [0, 4) takes two from
    each, block [4, 8) the other two; with a single thread the order is one
    shuffle of 0..7."""
    o = O.pgen_perm_order(8, 1, 4, 2, 112)
    assert set(o[:2].tolist()) | set(o[4:6].tolist()) == {0, 1, 2, 3}
    assert set(o[2:4].tolist()) | set(o[6:8].tolist()) == {4, 5, 6, 7}
    one = O.pgen_perm_order(8, 1, 4, 1, 112)
    np.testing.assert_array_equal(one, O.shuffle_parts(112, [np.arange(8)])[0])


@pytest.mark.parametrize("case", [c for c in sorted(p.name for p in GOLDEN.iterdir() if (p / "meta.json").exists())
                                  if "pgen" in load_meta(c)])
def test_golden_pgen_regenerates(case, golden_pgen, golden_inputs):
    """The .pgen every golden --pgen run read is written again bit-identically
    from the committed .bed, and decodes back to that .bed."""
    prefix = golden_pgen(case)
    rows, _ = _bed_rows(golden_inputs[load_meta(case)["input"]])
    hard = bool(load_meta(case)["pgen"].get("dosage"))
    np.testing.assert_array_equal(api.pgen_decode(prefix, hardcall=hard), rows)
