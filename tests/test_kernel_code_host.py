"""The tracking kernel's element code (xt_ops.h and the element-program
builder), compiled for the host, against the CPU oracle with the same
elementary-function library: identical IEEE values for every coordinate and
loss record.  The GPU tests repeat these comparisons with the device build."""
import numpy as np
import pytest

import xtrack_amd as xt
from helpers import mixed_line, random_particles, compare_exact, guard_edge_particles
from oracle_backend import OracleBackend, HostKernelBackend


def _run(line_fn, p, **kw):
    lh = line_fn()
    lh.build_tracker(backend=HostKernelBackend())
    lc = line_fn()
    lc.build_tracker(backend=OracleBackend())
    ph, pc = p.copy(), p.copy()
    lh.track(ph, **kw)
    lc.track(pc, **kw)
    return lh, lc, ph, pc


@pytest.mark.parametrize('n_turns', [1, 15])
def test_mixed_ring(n_turns):
    p = random_particles(800, seed=n_turns, scale=4e-3)
    _, _, ph, pc = _run(lambda: mixed_line(n_cells=3), p, num_turns=n_turns)
    assert not compare_exact(ph, pc)


def test_ions_chi_not_one():
    rng = np.random.default_rng(11)
    n = 400
    p = xt.Particles(p0c=7e9, x=rng.normal(0, 2e-3, n), px=rng.normal(0, 2e-4, n),
                     y=rng.normal(0, 2e-3, n), py=rng.normal(0, 2e-4, n),
                     zeta=rng.normal(0, 0.05, n), delta=rng.normal(0, 1e-3, n),
                     charge_ratio=rng.choice([0.5, 1.0, 2.0], n),
                     mass_ratio=rng.choice([1.0, 2.0, 4.0], n))
    _, _, ph, pc = _run(lambda: mixed_line(n_cells=2), p, num_turns=6)
    assert not compare_exact(ph, pc)


def test_nonfinite_and_extreme():
    rng = np.random.default_rng(12)
    n = 400
    x = rng.normal(0, 2e-3, n)
    px = rng.normal(0, 2e-4, n)
    x[:40] = np.nan
    px[40:80] = np.inf
    px[80:120] = rng.uniform(0.5, 3.0, 40)
    x[120:160] = rng.uniform(-0.99, 0.99, 40)
    p = xt.Particles(p0c=7e9, x=x, px=px, y=rng.normal(0, 1e-3, n), delta=rng.normal(0, 1e-3, n))
    _, _, ph, pc = _run(lambda: mixed_line(n_cells=2, aperture=0.5), p, num_turns=3)
    assert not compare_exact(ph, pc)
    assert set(np.unique(pc.sorted_by_id()['state'])) >= {-1, 0}


def test_reciprocal_division_range_edges():
    """Lanes on both sides of the reciprocal-division ranges (px = +-0,
    subnormal, ~2^-900; pz < 2^-16; delta near -1) take the literal form
    where needed and match the oracle bit for bit."""
    p = guard_edge_particles(1200, seed=4)
    _, _, ph, pc = _run(lambda: mixed_line(n_cells=2, aperture=0.95), p, num_turns=2)
    assert not compare_exact(ph, pc)
    p = guard_edge_particles(1200, seed=5)
    _, _, ph, pc = _run(lambda: mixed_line(n_cells=2, aperture=0.95), p, num_elements=7)
    assert not compare_exact(ph, pc)


@pytest.mark.parametrize('mode', [True, 'ONE_TURN_EBE'])
def test_monitors(mode):
    p = random_particles(200, seed=5, scale=3e-3)
    nt = 5 if mode is True else 1
    lh, lc, ph, pc = _run(lambda: mixed_line(n_cells=2), p, num_turns=nt,
                          turn_by_turn_monitor=mode)
    for ff in ('x', 'px', 'y', 'py', 'zeta', 'delta', 'at_turn', 'at_element', 'state'):
        assert np.array_equal(getattr(lh.record_last_track, ff),
                              getattr(lc.record_last_track, ff), equal_nan=True), ff


def test_thin_fodo_and_collider_cells():
    from xtrack_amd import lattices
    line_fn = lattices.thin_fodo_ring
    p = random_particles(300, seed=2, scale=2e-3, p0c=26e9)
    _, _, ph, pc = _run(line_fn, p, num_turns=20)
    assert not compare_exact(ph, pc)
    p = random_particles(64, seed=3, scale=5e-4, p0c=7e12, spread=0.3)
    _, _, ph, pc = _run(lambda: lattices.collider_ring(n_arcs=1, n_cells=2, k2f=0.03,
                                                       k2d=-0.06), p, num_turns=3)
    assert not compare_exact(ph, pc)


def test_reciprocal_division_is_correctly_rounded():
    """div_rcp(a, b, RN(1/b)) == a/b bit for bit in a range wider than the
    kernel uses (dividends within [2^-948, 2^948], |b| in [2^-16, 2^16]):
    |a| in [2^-960, 2^960], including
    quotients constructed to lie next to rounding midpoints."""
    import ctypes
    from oracle_backend import host_kernel_lib
    lib = host_kernel_lib()
    lib.xth_check_div_rcp.restype = ctypes.c_int64
    lib.xth_check_div_rcp.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]
    rng = np.random.default_rng(0)
    n = 2_000_000
    a = rng.uniform(-2, 2, n) * 2.0 ** rng.integers(-900, 900, n)
    b = rng.uniform(0.5, 2, n) * rng.choice([-1, 1], n) * 2.0 ** rng.integers(-16, 16, n)
    assert lib.xth_check_div_rcp(a.ctypes.data, b.ctypes.data, n) == 0
    # near-midpoint quotients: a = (q + ulp/2 * (1 + tiny)) * b with q a double
    q = rng.uniform(1, 2, n)
    b2 = rng.uniform(1, 2, n)
    a2 = (q + np.spacing(q) * 0.5) * b2
    assert lib.xth_check_div_rcp(a2.ctypes.data, b2.ctypes.data, n) == 0
    # divisors with all-ones significands (reciprocal next to a power of 2)
    b3 = np.nextafter(2.0 ** rng.integers(-10, 10, n).astype(float), 0)
    a3 = rng.uniform(-4, 4, n)
    assert lib.xth_check_div_rcp(a3.ctypes.data, b3.ctypes.data, n) == 0


@pytest.mark.parametrize('kw', [dict(num_turns=2), dict(ele_start=3, ele_stop=11),
                                dict(num_turns=1, turn_by_turn_monitor='ONE_TURN_EBE')])
def test_backtracking(kw):
    """Backtracking programs and the reversed element loop against the
    oracle, after a forward pass; losses included."""
    p = random_particles(400, seed=8, scale=5e-3)
    lh, lc, ph, pc = _run(lambda: mixed_line(n_cells=2), p, num_turns=2)
    lh.track(ph, backtrack=True, **kw)
    lc.track(pc, backtrack=True, **kw)
    assert not compare_exact(ph, pc)
    if 'turn_by_turn_monitor' in kw:
        for ff in ('x', 'px', 'zeta', 'at_element', 'at_turn', 'state'):
            assert np.array_equal(getattr(lh.record_last_track, ff),
                                  getattr(lc.record_last_track, ff), equal_nan=True), ff
