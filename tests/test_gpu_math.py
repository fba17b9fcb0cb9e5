"""Device elementary functions, sqrt and division against the host build
of the same library: bit-identical for every input."""
import ctypes

import numpy as np
import pytest

from oracle_backend import oracle_math

pytestmark = pytest.mark.gpu

FN = {'sin': 0, 'cos': 1, 'sinh': 2, 'cosh': 3, 'exp': 4, 'tan': 5, 'sqrt': 6, 'div': 7}


def _inputs(fn, rng, n):
    if fn in ('sinh', 'cosh', 'exp'):
        return np.concatenate([rng.uniform(-2, 2, n), rng.uniform(-720, 720, n),
                               [0.0, -0.0, 1.0, -1.0, 709.5, -745.0, np.inf, -np.inf, np.nan]])
    if fn == 'sqrt':
        return np.concatenate([10 ** rng.uniform(-300, 300, n), rng.uniform(0, 4, n),
                               [0.0, -0.0, 1.0, 4.0, np.inf, 5e-324, 2.2250738585072014e-308]])
    return np.concatenate([rng.uniform(-1, 1, n), rng.uniform(-100, 100, n),
                           rng.uniform(-1e6, 1e6, n), 10 ** rng.uniform(8, 300, n // 4),
                           [0.0, -0.0, np.pi / 4, np.pi / 2, 1e22, np.inf, np.nan]])


@pytest.mark.parametrize('fn', list(FN))
def test_device_math_bit_exact(fn):
    import torch
    from xtrack_amd import _native
    lib = _native.load()
    rng = np.random.default_rng(FN[fn])
    n = 200000
    if fn == 'div':
        a = rng.normal(size=n) * 10 ** rng.uniform(-30, 30, n)
        b = rng.normal(size=n) * 10 ** rng.uniform(-30, 30, n)
        x = np.concatenate([a, b])
        m = n
    else:
        x = _inputs(fn, rng, n)
        m = len(x)
    xd = torch.from_numpy(x).cuda()
    yd = torch.empty(m, dtype=torch.float64, device='cuda')
    _native.check(lib.xtm_math_eval(FN[fn], ctypes.c_void_p(xd.data_ptr()),
                                    ctypes.c_void_p(yd.data_ptr()), m, None))
    torch.cuda.synchronize()
    yg = yd.cpu().numpy()
    yc = oracle_math(FN[fn], x, 'xm')
    same = (yg.view(np.int64) == yc.view(np.int64)) | (np.isnan(yg) & np.isnan(yc))
    assert same.all(), (fn, x[:m][~same][:5], yg[~same][:5], yc[~same][:5])
