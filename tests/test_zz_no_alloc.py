"""pcaone_packed_gemm_noalloc: the packed product on device pointers with
every intermediate in the library's fixed device scratch array.  Its output
must equal pcaone_packed_gemm's (the allocating entry point, same kernels)
bit for bit, on two calls in a row on one context, the second smaller (the
scratch keeps the first call's contents), for every single-form precision."""
import ctypes as C

import numpy as np
import pytest
import torch

from pcaone_amd import _lib, api

pytestmark = pytest.mark.gpu


def _noalloc(ctx, precision, packed, ncodes, W):
    dev = torch.device("cuda:0")
    rows = torch.from_numpy(np.ascontiguousarray(packed)).to(dev)
    Wd = torch.from_numpy(np.asfortranarray(W).ravel(order="F").copy()).to(dev)
    Y = torch.full((W.shape[1] * packed.shape[0],), float("nan"), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    st = _lib.lib().pcaone_packed_gemm_noalloc(ctx._h, api.PRECISIONS[precision], C.c_void_p(rows.data_ptr()),
                                                packed.shape[0], ncodes, C.c_void_p(Wd.data_ptr()), W.shape[1],
                                                C.c_void_p(Y.data_ptr()))
    _lib.check(st, ctx._h)
    torch.cuda.synchronize()
    return Y.cpu().numpy().reshape(W.shape[1], packed.shape[0]).T


@pytest.mark.parametrize("precision", ["fp64", "fp32", "bf16", "int8", "fp8"])
def test_noalloc_matches_allocating_entry_point(precision):
    rng = np.random.default_rng(11)
    with api.Context(0) as ctx:
        for nrows, ncodes, L in ((700, 3000, 80), (129, 515, 20)):  # the second call is smaller
            print(f"noalloc {precision} nrows={nrows} ncodes={ncodes} L={L}", flush=True)
            packed = rng.integers(0, 256, size=(nrows, (ncodes + 3) // 4), dtype=np.uint8)
            W = rng.standard_normal((ncodes, L))
            ref = ctx.packed_gemm(precision, packed, ncodes, W)
            out = _noalloc(ctx, precision, packed, ncodes, W)
            np.testing.assert_array_equal(out, ref)


def test_noalloc_rejects_a_workspace_beyond_the_scratch():
    with api.Context(0) as ctx:
        nrows, ncodes, L = 65536, 65536, 16  # 64 MiB of packed codes alone
        packed = np.zeros((1, 1), np.uint8)
        st = _lib.lib().pcaone_packed_gemm_noalloc(ctx._h, api.PRECISIONS["fp32"], C.c_void_p(16), nrows, ncodes,
                                                    C.c_void_p(16), L, C.c_void_p(16))
        assert st == 1 and b"exceeds" in _lib.lib().pcaone_last_error(ctx._h)
