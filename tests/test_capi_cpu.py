"""CPU-side checks of libpcaone_amd.so: exported symbols, host FP64 helpers,
parameter defaults and the no-device error path (no GPU compute here)."""
import ctypes as C
import os
import subprocess
import re
from pathlib import Path

import numpy as np
import pytest

from pcaone_amd import _lib
from pcaone_amd import api

HEADER = Path(__file__).resolve().parents[1] / "include" / "pcaone_amd.h"


def header_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pcaone_\w+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    L = _lib.lib()
    names = header_functions()
    assert len(names) >= 30
    for n in names:
        assert hasattr(L, n), f"{n} declared in include/pcaone_amd.h but not exported"


def test_binding_covers_header():
    assert set(header_functions()) == set(_lib._SIGS)


STRUCTS = [("pcaone_params", _lib.Params), ("pcaone_run_info", _lib.RunInfo),
           ("pcaone_synth_spec", _lib.SynthSpec), ("pcaone_batch_info", _lib.BatchInfo)]


def test_ctypes_structs_match_header_layout(tmp_path):
    """Every ctypes structure has the size and field offsets that a C compiler
    gives the header's struct (a caller's buffer of the wrong size would be
    overrun by pcaone_params_default and the run-info writers)."""
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "pcaone_amd.h"', "int main(void) {"]
    for cname, py in STRUCTS:
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'  printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", f"-I{HEADER.parent}", str(src), "-o", str(exe)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines():
        cname, key, val = line.split()
        got[(cname, key)] = int(val)
    for cname, py in STRUCTS:
        assert got[(cname, "size")] == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[(cname, f)] == getattr(py, f).offset, f"{cname}.{f}"


def test_params_default_match_reference():
    p = api.default_params()
    # src/Cmd.hpp defaults and src/Cmd.cpp:41-45 (-d defaults to 2)
    assert (p.svd, p.k, p.oversamples, p.maxp, p.bands, p.maxiter) == (2, 10, 10, 20, 64, 100)
    assert (p.tol, p.tolem, p.seed, p.rand, p.ploidy) == (1e-4, 1e-5, 112, 1, 2)


def test_open_without_device_reports_status():
    if api.device_count() > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(_lib.PcaoneError) as e:
        api.Context(0)
    assert e.value.status == _lib.ERR_NO_DEVICE
    assert "no HIP device" in str(e.value)


def _host(fn, *arrays, n):
    outs = [np.zeros(a, dtype=np.float64) for a in arrays]
    return outs


def test_host_sym_eig_matches_lapack():
    rng = np.random.default_rng(5)
    for n in (1, 2, 7, 20, 80, 128):
        A = rng.standard_normal((n, n))
        A = A @ A.T + np.diag(rng.uniform(0, 1, n))
        w = np.empty(n)
        V = np.empty((n, n))
        _lib.check(_lib.lib().pcaone_host_sym_eig(n, np.ascontiguousarray(A).ctypes.data, w.ctypes.data,
                                                  V.ctypes.data))
        wref = np.sort(np.linalg.eigvalsh(A))[::-1]
        np.testing.assert_allclose(w, wref, rtol=1e-12, atol=1e-12 * wref[0])
        np.testing.assert_allclose(V @ np.diag(w) @ V.T, A, atol=1e-11 * np.abs(A).max())
        np.testing.assert_allclose(V.T @ V, np.eye(n), atol=1e-12)


def test_host_householder_signs_reproduce_lapack_q():
    rng = np.random.default_rng(6)
    for trial in range(200):
        n = int(rng.integers(30, 200))
        L = int(rng.integers(1, 30))
        A = rng.standard_normal((n, L)) * rng.uniform(0.1, 10, L)
        Qh = np.linalg.qr(A)[0]  # Householder, beta = -sign(alpha) ||x||
        R = np.linalg.cholesky(A.T @ A).T
        Qc = A @ np.linalg.inv(R)
        s = np.empty(L)
        Y = np.ascontiguousarray(Qc[:L, :L])
        _lib.check(_lib.lib().pcaone_host_householder_signs(L, Y.ctypes.data, s.ctypes.data))
        np.testing.assert_allclose(Qc * s, Qh, atol=1e-9)


def test_host_cholesky_and_breakdown():
    rng = np.random.default_rng(7)
    A = rng.standard_normal((12, 12))
    A = A @ A.T + np.eye(12)
    R = np.empty((12, 12))
    _lib.check(_lib.lib().pcaone_host_chol_upper(12, A.ctypes.data, R.ctypes.data))
    np.testing.assert_allclose(R.T @ R, A, rtol=1e-12, atol=1e-12)
    assert np.allclose(np.tril(R, -1), 0) and np.all(np.diag(R) > 0)
    B = -np.eye(3)
    st = _lib.lib().pcaone_host_chol_upper(3, B.ctypes.data, R.ctypes.data)
    assert st == 5  # PCAONE_ERR_NUMERIC
    assert b"Cholesky" in _lib.lib().pcaone_last_error(None)


@pytest.mark.parametrize("n,L,rank", [(50, 7, 7), (300, 16, 11), (2000, 20, 20), (64, 12, 1)])
def test_host_householder_q_is_lapacks(n, L, rank):
    """The Householder thin Q (the fallback for a near rank-deficient H) equals
    LAPACK's dgeqrf/dorgqr Q (numpy) column for column, signs included, for
    full-rank X; for rank-deficient X its first `rank` columns equal LAPACK's
    and all L columns are orthonormal."""
    rng = np.random.default_rng(n + L)
    X = rng.standard_normal((n, rank)) @ rng.standard_normal((rank, L))
    Q = np.ascontiguousarray(X.copy())
    _lib.check(_lib.lib().pcaone_host_householder_q(n, L, Q.ctypes.data))
    np.testing.assert_allclose(Q.T @ Q, np.eye(L), atol=1e-13)
    Qref = np.linalg.qr(X, mode="reduced")[0]
    np.testing.assert_allclose(Q[:, :rank], Qref[:, :rank], atol=1e-12)
    if rank == L:
        np.testing.assert_allclose(Q, Qref, atol=1e-12)


def test_null_arguments_are_rejected():
    L = _lib.lib()
    assert L.pcaone_synth_packed(None, None, None) == 1
    assert L.pcaone_close(None) == 0
    assert L.pcaone_sketch(None, 0, 1, 0) == 1


@pytest.mark.parametrize("seed,gaussian,n,l", [(112, 1, 2000, 20), (7, 0, 333, 5), (112, 1, 50000, 8),
                                               (0, 1, 101, 3), (-3, 1, 77, 9), (2**31 - 1, 0, 10, 4)])
def test_parallel_omega_is_the_libstdcxx_draw(seed, gaussian, n, l):
    from oracle import pcaone_oracle as O
    out = np.empty((l, n))  # column-major n x l
    _lib.check(_lib.lib().pcaone_host_omega(seed, gaussian, n, l, out.ctypes.data))
    ref = O.omega(seed, bool(gaussian), n, l)
    assert np.array_equal(out.T, ref)


# ---------------------------------------------------------------- out-of-core plan and shuffle order
PLAN_GRID = [
    # (nsamples, nsnps, svd, k, oversamples, bands, memory)
    (200, 2000, 1, 5, 10, 64, 0.002), (200, 2000, 1, 5, 10, 64, 0.0003), (200, 2000, 2, 5, 10, 16, 0.00068),
    (200, 2000, 2, 5, 10, 16, 0.0007), (300, 3000, 1, 3, 10, 64, 0.002), (300, 3000, 2, 3, 10, 16, 0.0009),
    (300, 3000, 2, 3, 10, 16, 0.002), (50000, 1000000, 1, 40, 20, 64, 16.0), (50000, 1000000, 2, 40, 20, 64, 1.0),
    (10000, 500000, 2, 20, 10, 64, 2.0), (2000, 20000, 1, 10, 10, 64, 0.05), (20000, 300000, 1, 10, 10, 64, 0.4),
]


def _library_plan(n, m, svd, k, os, bands, memory):
    p = api.default_params(svd=svd, k=k, oversamples=os, bands=bands, memory=memory)
    bs, nb, bf = C.c_uint64(), C.c_uint64(), C.c_uint64()
    start = np.zeros(1 << 16, dtype=np.uint64)
    rc = _lib.lib().pcaone_block_plan(C.byref(p), n, m, C.byref(bs), C.byref(nb), C.byref(bf), start.ctypes.data,
                                      start.size)
    return rc, bs.value, nb.value, bf.value, start[: nb.value]


@pytest.mark.parametrize("n,m,svd,k,os,bands,memory", PLAN_GRID)
def test_block_plan_library_matches_oracle(n, m, svd, k, os, bands, memory):
    from oracle import pcaone_oracle as O
    rc, bs, nb, bf, start = _library_plan(n, m, svd, k, os, bands, memory)
    try:
        ref = O.block_plan(O.Params(svd=svd, k=k, oversamples=os, bands=bands, memory=memory), n, m)
    except ValueError:
        assert rc == 1  # the reference refuses a plan with a single block
        return
    assert rc == 0
    assert (bs, nb, bf) == (ref.blocksize, ref.nblocks, ref.band_factor)
    assert list(start) == ref.start


def test_block_plan_single_block_is_an_error():
    rc, *_ = _library_plan(200, 2000, 1, 5, 10, 64, 1.0)  # 1 GB holds all 2,000 SNPs
    assert rc == 1
    assert b"only one block" in _lib.lib().pcaone_last_error(None)


@pytest.mark.parametrize("nsamples,nsnps,gb,bands", [
    (200, 2000, 2, 16), (300, 3000, 2, 16), (300, 3001, 2, 64), (777, 64, 2, 64), (5, 1000, 2, 4),
    # nominal sample counts that make the read buffer hold fewer SNPs than the file: several buffers
    (40_000_000, 1000, 1, 4), (40_000_000, 1003, 1, 4), (400_000_000, 997, 3, 8), (100_000_000, 4099, 2, 16),
])
def test_permute_plink_order_library_matches_oracle(nsamples, nsnps, gb, bands):
    from oracle import pcaone_oracle as O
    out = np.empty(nsnps, dtype=np.int32)
    _lib.check(_lib.lib().pcaone_permute_plink_order(nsamples, nsnps, gb, bands, out.ctypes.data))
    ref = O.permute_plink_order(nsamples, nsnps, gb, bands)
    assert np.array_equal(out, ref)
    assert np.array_equal(np.sort(out), np.arange(nsnps))  # a permutation


def test_permute_plink_order_single_buffer_is_band_interleave():
    """One buffer and nsnps a multiple of the bands: band b's run holds SNPs b, b + w, b + 2w, ..."""
    n, m, w = 200, 2000, 16
    out = np.empty(m, dtype=np.int32)
    _lib.check(_lib.lib().pcaone_permute_plink_order(n, m, 2, w, out.ctypes.data))
    assert np.array_equal(out, np.arange(m).reshape(m // w, w).T.ravel())


# ---------------------------------------------------------------- FileBed constructor (no device)
def _bed_dims(prefix):
    n, m = C.c_uint32(), C.c_uint64()
    st = _lib.lib().pcaone_bed_dims(str(prefix).encode(), C.byref(n), C.byref(m))
    return st, n.value, m.value, _lib.lib().pcaone_last_error(None).decode()


def _write_set(prefix, n, m, header=b"\x6c\x1b\x01"):
    Path(str(prefix) + ".fam").write_text("".join(f"f{i} i{i} 0 0 0 -9\n" for i in range(n)))
    Path(str(prefix) + ".bim").write_text("".join(f"1 s{j} 0 {j + 1} A G\n" for j in range(m)))
    Path(str(prefix) + ".bed").write_bytes(header + bytes(((n + 3) // 4) * m))


def test_bed_dims_counts_fam_and_bim_lines(tmp_path):
    """src/FilePlink.hpp:14-17: N = lines of .fam, M = lines of .bim."""
    _write_set(tmp_path / "ok", 13, 7)
    st, n, m, _ = _bed_dims(tmp_path / "ok")
    assert (st, n, m) == (0, 13, 7)


@pytest.mark.parametrize("header", [b"\x6c\x1b\x00", b"\x00\x1b\x01", b"\x6c\x00\x01", b"\x6c\x1b", b""])
def test_bed_bad_magic_is_an_io_error(tmp_path, header):
    """src/FilePlink.hpp:24-27: a .bed whose first three bytes are not 6c 1b 01
    (SNP-major PLINK) is rejected with the reference's message."""
    _write_set(tmp_path / "bad", 8, 3, header=header)
    st, _, _, msg = _bed_dims(tmp_path / "bad")
    assert st == 4
    assert msg == "Incorrect magic number in plink bed file."


def test_bed_missing_file_is_an_io_error(tmp_path):
    """src/FilePlink.hpp:21-22: a missing .bed cannot be opened."""
    _write_set(tmp_path / "nobed", 8, 3)
    Path(str(tmp_path / "nobed") + ".bed").unlink()
    st, _, _, msg = _bed_dims(tmp_path / "nobed")
    assert (st, msg) == (4, "Cannot open bed file.")
    st, _, _, msg = _bed_dims(tmp_path / "absent")
    assert st == 4 and "can not open" in msg


REF_SRC = Path(os.environ.get("PCAONE_REF", "/root/reference"))


@pytest.mark.skipif(not (REF_SRC / "src" / "Halko.hpp").exists(), reason="reference sources absent")
def test_binding_compiles_against_reference_headers():
    """The maintainer-side binding (AmdFileBed : Data, AmdNormalRsvdOpData :
    NormalRsvdOpData, AmdFancyRsvdOpData : FancyRsvdOpData, the level-1 and
    level-2 drivers) compiles against the reference's own src/Halko.hpp and
    src/Data.hpp with the reference's compiler flags."""
    root = Path(__file__).resolve().parents[1]
    cmd = ["g++", "-fsyntax-only", "-std=c++17", "-fopenmp", "-DVERSION=\"0.7.2\"", "-w",
           f"-I{REF_SRC / 'src'}", f"-I{REF_SRC / 'external'}", f"-I{REF_SRC / 'external' / 'zstd' / 'lib'}",
           f"-I{root / 'include'}", f"-I{root / 'pcaone_amd' / 'binding'}", str(root / "oracle" / "ref_driver_amd.cpp")]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr[-3000:]
