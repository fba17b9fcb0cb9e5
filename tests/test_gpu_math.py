"""Device elementary functions, sqrt and division against the host build
of the same library: bit-identical for every input."""
import ctypes

import numpy as np
import pytest

from oracle_backend import oracle_math

pytestmark = pytest.mark.gpu

FN = {'sin': 0, 'cos': 1, 'sinh': 2, 'cosh': 3, 'exp': 4, 'tan': 5, 'sqrt': 6, 'div': 7,
      'atan2': 11, 'asin': 12}


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
    if fn in ('div', 'atan2'):
        a = rng.normal(size=n) * 10 ** rng.uniform(-30, 30, n)
        b = rng.normal(size=n) * 10 ** rng.uniform(-30, 30, n)
        if fn == 'atan2':
            sp = np.array([0.0, -0.0, 1.0, -1.0, np.inf, -np.inf, np.nan, 1e-300])
            a[:64], b[:64] = np.repeat(sp, 8), np.tile(sp, 8)
        x = np.concatenate([a, b])
        m = n
    elif fn == 'asin':
        x = np.concatenate([rng.uniform(-1, 1, n), rng.uniform(-1e-8, 1e-8, n // 4),
                            1 - 10 ** rng.uniform(-16, -1, n // 4),
                            [0.0, -0.0, 1.0, -1.0, 0.5, 1.0 + 2e-16, np.nan]])
        m = len(x)
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


def _eval(fn, x, m):
    import torch
    from xtrack_amd import _native
    lib = _native.load()
    xd = torch.from_numpy(np.ascontiguousarray(x)).cuda()
    yd = torch.empty(m, dtype=torch.float64, device='cuda')
    _native.check(lib.xtm_math_eval(fn, ctypes.c_void_p(xd.data_ptr()),
                                    ctypes.c_void_p(yd.data_ptr()), m, None))
    torch.cuda.synchronize()
    return yd.cpu().numpy()


def _same_bits(a, b):
    return a.view(np.int64) == b.view(np.int64)


def test_in_range_reciprocal_and_quotient_are_ieee():
    """The kernel's shortened reciprocal and division sequences equal the
    IEEE quotient (numpy on the host) over their whole operand ranges,
    including all-ones significands, powers of two and quotients next to
    rounding midpoints."""
    rng = np.random.default_rng(8)
    n = 2_000_000
    sgn = lambda k: rng.choice([-1.0, 1.0], k)
    b = np.concatenate([
        rng.uniform(1, 2, n) * sgn(n) * 2.0 ** rng.integers(-16, 16, n),
        np.nextafter(2.0 ** rng.integers(-15, 17, n // 8).astype(float), 0) * sgn(n // 8),
        2.0 ** rng.integers(-16, 17, 64).astype(float),
        [2.0 ** -16, -(2.0 ** -16), 2.0 ** 16, 1.0, -1.0]])
    y = _eval(8, b, len(b))
    ok = _same_bits(y, 1.0 / b)
    assert ok.all(), (b[~ok][:5], y[~ok][:5])
    m = len(b)
    a = np.concatenate([rng.uniform(1, 2, m - n // 4) * sgn(m - n // 4)
                        * 2.0 ** rng.integers(-940, 940, m - n // 4),
                        (lambda q: (q + np.spacing(q) * 0.5))(rng.uniform(1, 2, n // 4))])
    bb = b.copy()
    bb[m - n // 4:] = rng.uniform(1, 2, n // 4)
    a[m - n // 4:] *= bb[m - n // 4:]
    y = _eval(10, np.concatenate([a, bb]), m)
    ok = _same_bits(y, a / bb)
    assert ok.all(), (a[~ok][:5], bb[~ok][:5], y[~ok][:5])


def test_in_range_sqrt_is_ieee():
    rng = np.random.default_rng(9)
    n = 2_000_000
    r = rng.uniform(2.0 ** -16, 2.0 ** 16, n)
    x = np.concatenate([
        rng.uniform(1, 2, n) * 2.0 ** rng.integers(-32, 32, n),
        r * r, np.nextafter(r * r, 0), np.nextafter(r * r, np.inf),
        (lambda s: s * s + np.spacing(s * s) * 0.5)(rng.uniform(1, 2, n // 4)),
        [2.0 ** -32, 2.0 ** 32, 1.0, 2.0, 0.5]])
    x = x[(x >= 2.0 ** -32) & (x <= 2.0 ** 32)]
    y = _eval(9, x, len(x))
    ok = _same_bits(y, np.sqrt(x))
    assert ok.all(), (x[~ok][:5], y[~ok][:5])
