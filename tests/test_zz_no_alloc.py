"""The allocation-free transform (semx_dft_r_noalloc) against semx_dft_r,
through the C ABI.

semx_dft_r_noalloc works in the library's static device scratch; its
transform matrix is formed on the device in FP64 (semx_dft_r forms it on the
host in long double), so the two agree to rounding: bar rel L2 <= 1e-14 per
transform, FP64.  Two calls in a row, the second with a smaller input, so
that the second runs on the array the first left behind.
"""
import ctypes as C

import numpy as np
import pytest

import semtex_amd as sx

pytestmark = pytest.mark.gpu

SEMX_E_SCRATCH = 10


def _run(L, torch, fn, x, tlen, ntrn, sign):
    t = torch.from_numpy(x.ravel().copy()).to("cuda:0")
    st = getattr(L, fn)(C.c_void_p(t.data_ptr()), tlen, ntrn, sign,
                        C.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert st == 0, (fn, st, L.semx_error_string(st))
    torch.cuda.synchronize()
    return t.cpu().numpy().reshape(tlen, ntrn)


@pytest.mark.parametrize("sign", [1, -1])
def test_noalloc_matches_dft_r_two_calls(cuda, sign):
    import torch
    L = sx.load()
    assert L.semx_device_init(0) == 0
    # the DNS field (96 planes of 24x24 elements at np=9), then a smaller one
    for tlen, ntrn, seed in ((96, 46656, 1), (64, 2 * 1237, 2)):
        print(f"[noalloc] sign={sign} tlen={tlen} ntrn={ntrn}", flush=True)
        x = np.random.default_rng(seed).standard_normal((tlen, ntrn))
        ref = _run(L, torch, "semx_dft_r", x, tlen, ntrn, sign)
        got = _run(L, torch, "semx_dft_r_noalloc", x, tlen, ntrn, sign)
        err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        assert err <= 1e-14, (tlen, ntrn, err)


@pytest.mark.parametrize("tlen,ntrn", [(6, 10), (160, 2 * 331), (256, 2 * 999)])
def test_noalloc_other_lengths(cuda, tlen, ntrn):
    """Lengths served by the vector-FMA kernel (> 128) and short ones."""
    import torch
    L = sx.load()
    print(f"[noalloc] tlen={tlen} ntrn={ntrn}", flush=True)
    x = np.random.default_rng(tlen).standard_normal((tlen, ntrn))
    for sign in (1, -1):
        ref = _run(L, torch, "semx_dft_r", x, tlen, ntrn, sign)
        got = _run(L, torch, "semx_dft_r_noalloc", x, tlen, ntrn, sign)
        assert np.linalg.norm(got - ref) <= 1e-14 * np.linalg.norm(ref)


def test_noalloc_rejects_oversize(cuda):
    """Sizes beyond the static scratch are refused before any device work."""
    import torch
    L = sx.load()
    t = torch.zeros(8, dtype=torch.float64, device="cuda:0")
    p = C.c_void_p(t.data_ptr())
    big = (6 << 20) // 96 + 2          # even, 96 * big > SEMX_DFT_NOALLOC_MAX
    assert L.semx_dft_r_noalloc(p, 96, big, 1, None) == SEMX_E_SCRATCH
    assert L.semx_dft_r_noalloc(p, 270, 2, 1, None) == SEMX_E_SCRATCH
    assert L.semx_dft_r_noalloc(p, 7, 2, 1, None) == 1        # dDFTr's length check
