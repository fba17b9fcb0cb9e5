"""xtm_gemm_bf16_noalloc (the entry point that allocates nothing: its
k-split partials live in a file-scope device array) against
xtm_gemm_bf16 (library-allocated workspace), through the C ABI: identical
FP32 and BF16 results, on calls in a row -- each smaller than the one
before, so the scratch array holds the previous call's partials when the
next starts -- and invalid arguments refused with a message."""
import ctypes

import numpy as np
import pytest

from xtrack_amd import _native

pytestmark = pytest.mark.gpu


def _call(fn, a, b, M, N, K, ta, tb, bias, relu):
    import torch
    lib = _native.load()
    sam, sak = (1, M) if ta else (K, 1)
    sbk, sbn = (1, K) if tb else (N, 1)
    c = torch.empty((M, N), dtype=torch.float32, device=a.device)
    cb = torch.empty((M, N), dtype=torch.int16, device=a.device)
    stream = ctypes.c_void_p(torch.cuda.current_stream(a.device).cuda_stream)
    st = getattr(lib, fn)(ctypes.c_void_p(a.data_ptr()), sam, sak, ctypes.c_void_p(b.data_ptr()),
                          sbk, sbn, ctypes.c_void_p(c.data_ptr()), N, ctypes.c_void_p(cb.data_ptr()),
                          ctypes.c_void_p(bias.data_ptr()) if bias is not None else None,
                          int(relu), M, N, K, stream)
    return st, c, cb


def _operands(M, N, K, ta, tb, seed):
    import torch
    rng = np.random.default_rng(seed)
    a = rng.normal(size=(K, M) if ta else (M, K)).astype(np.float32)
    b = rng.normal(size=(N, K) if tb else (K, N)).astype(np.float32)
    dev = torch.device('cuda:0')
    to = lambda x: torch.from_numpy(x).to(dev).to(torch.bfloat16).view(torch.int16).contiguous()
    bias = torch.from_numpy(rng.normal(size=N).astype(np.float32)).to(dev)
    return to(a), to(b), bias


def test_noalloc_gemm_equals_allocating_gemm():
    import torch
    print('[no-alloc gemm]', flush=True)
    lib = _native.load()
    assert lib.xtm_gemm_bf16_scratch_floats() == 1 << 20
    # two calls in a row, the second smaller; both split k over blocks
    for (M, N, K, ta, tb, relu), seed in (((256, 256, 4096, True, False, False), 1),
                                          ((64, 128, 2000, False, True, True), 2),
                                          ((70, 33, 17, False, False, True), 3)):
        a, b, bias = _operands(M, N, K, ta, tb, seed)
        st1, c1, cb1 = _call('xtm_gemm_bf16_noalloc', a, b, M, N, K, ta, tb, bias, relu)
        st0, c0, cb0 = _call('xtm_gemm_bf16', a, b, M, N, K, ta, tb, bias, relu)
        torch.cuda.synchronize()
        assert st1 == 0 and st0 == 0
        assert torch.equal(c1.view(torch.int32), c0.view(torch.int32)), (M, N, K)
        assert torch.equal(cb1, cb0), (M, N, K)
        assert float(c0.abs().max()) > 0


def test_noalloc_gemm_argument_errors():
    """An invalid call is refused with a message (ldc < N here); the split
    chosen for any valid product keeps splits x M x N within the array, so
    its partials always fit."""
    lib = _native.load()
    import torch
    a, b, _ = _operands(128, 128, 64, False, False, 4)
    c = torch.empty((128, 128), dtype=torch.float32, device='cuda:0')
    # ldc < N: argument error with a message
    st = lib.xtm_gemm_bf16_noalloc(ctypes.c_void_p(a.data_ptr()), 64, 1,
                                   ctypes.c_void_p(b.data_ptr()), 128, 1,
                                   ctypes.c_void_p(c.data_ptr()), 100, None, None, 0, 128, 128, 64,
                                   None)
    assert st == _native.XTM_ERR_ARG
    assert b'ldc' in lib.xtm_last_error()
