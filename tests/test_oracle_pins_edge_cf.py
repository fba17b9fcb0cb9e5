"""Pins of two oracle maps that the whole-ring tests only check oracle
against kernel: the expanded combined-function dipole map with k0 != 0
(track_magnet_drift.h:91-213, the 'mat-kick-mat' core) and the linear
dipole edge kick (dipoleedge.h:17-50 -> track_dipole_edge_linear.h:30-40).

* Combined function: the reference's own test
  (tests/test_elements_thick.py:129-177) tracks the same bend with the
  expanded core and 100 kicks and with the exact rot-kick-rot core and 100
  kicks and requires agreement within 5e-9 (x, y), 2e-9 (px, py) and 1e-11
  (zeta, ptau); then the expanded core backtracks to the initial
  coordinates within 1e-11.  A dropped or wrongly signed term of the
  expanded map (the (k0 - h) offset, the path-length sums) moves x or zeta
  by far more than those tolerances (k0 - h = 1e-4 m^-1 shifts x by 5e-5 m
  over the metre), which the rot-kick-rot integrator, a different
  restatement (polar drifts and kicks), does not share.
* Linear edge: the kick is px += chi r21 (1 + delta_taper) x,
  py += chi r43 (1 + delta_taper) y with r21 = k tan(e1) and
  r43 = -k tan(e1 + e1_fd - 2 k hgap fint / cos(e1 + e1_fd)
  (1 + sin^2(e1 + e1_fd))) (dipole_edge.py:123-131), evaluated here in
  mpmath at 50 digits; backtracking flips the signs of r21 and r43.
  A transposed pair (r21 on y), a dropped chi or taper factor or a missing
  sign flip fails the closed form."""
import mpmath as mp
import numpy as np
import pytest

import xtrack_amd as xt
from xtrack_amd.constants import PROTON_MASS_EV
from oracle_backend import OracleBackend


def _p0():
    # tests/test_elements_thick.py:131-141
    return xt.Particles(mass0=PROTON_MASS_EV, beta0=0.5, x=0.003, px=0.001, y=-0.005,
                        py=0.001, zeta=0.1, delta=[-1e-3, -0.05e-3, 0, 0.05e-3, 1e-3])


def _bend(model, **kw):
    return xt.Bend(k0=1e-3, angle=0.9e-3, length=1, k1=0.001, knl=[0, 0, 0.02],
                   model=model, **kw)


@pytest.mark.parametrize('math', ['xm', 'libm'])
def test_combined_function_expanded_vs_rot_kick_rot(math):
    """test_elements_thick.py:143-167"""
    be = OracleBackend(math=math)
    lt = xt.Line(elements=[_bend('mat-kick-mat', num_multipole_kicks=100)])
    lt.build_tracker(backend=be)
    lr = xt.Line(elements=[_bend('rot-kick-rot', num_multipole_kicks=100)])
    lr.build_tracker(backend=be)
    p_test = _p0()
    p_ref = _p0()
    lt.track(p_test)
    lr.track(p_ref)
    t = p_test.sorted_by_id()
    r = p_ref.sorted_by_id()
    assert np.all(t['state'] == 1) and np.all(r['state'] == 1)
    for ff, atol in (('x', 5e-9), ('px', 2e-9), ('y', 5e-9), ('py', 2e-9),
                     ('zeta', 1e-11), ('ptau', 1e-11)):
        err = np.max(np.abs(t[ff] - r[ff]))
        assert err <= atol, (ff, err)
    # the map does move the particles by much more than the tolerance
    p0 = _p0().sorted_by_id()
    assert np.max(np.abs(t['x'] - p0['x'])) > 5e-4
    assert np.max(np.abs(t['zeta'] - p0['zeta'])) > 1e-9


@pytest.mark.parametrize('math', ['xm', 'libm'])
def test_combined_function_expanded_backtrack(math):
    """test_elements_thick.py:169-177: the expanded core tracked forward and
    then backward returns to the initial coordinates within 1e-11."""
    be = OracleBackend(math=math)
    line = xt.Line(elements=[_bend('mat-kick-mat', num_multipole_kicks=100)])
    line.build_tracker(backend=be)
    p = _p0()
    line.track(p)
    line.track(p, backtrack=True)
    q = p.sorted_by_id()
    p0 = _p0().sorted_by_id()
    for ff in ('x', 'px', 'y', 'py', 'zeta', 'ptau'):
        err = np.max(np.abs(q[ff] - p0[ff]))
        assert err <= 1e-11, (ff, err)


def test_combined_function_planted_error_is_caught():
    """The comparison above is not blind to a wrong expanded map: the same
    expanded bend with k0 replaced by h (a dropped (k0 - h) term) differs
    from rot-kick-rot far beyond the tolerance."""
    be = OracleBackend(math='libm')
    lt = xt.Line(elements=[xt.Bend(k0=0.9e-3, angle=0.9e-3, length=1, k1=0.001,
                                   knl=[0, 0, 0.02], model='mat-kick-mat',
                                   num_multipole_kicks=100)])
    lt.build_tracker(backend=be)
    lr = xt.Line(elements=[_bend('rot-kick-rot', num_multipole_kicks=100)])
    lr.build_tracker(backend=be)
    p_test, p_ref = _p0(), _p0()
    lt.track(p_test)
    lr.track(p_ref)
    assert np.max(np.abs(p_test.sorted_by_id()['x'] - p_ref.sorted_by_id()['x'])) > 1e-6


def _edge_coeffs(k, e1, e1_fd, hgap, fint):
    mp.mp.dps = 50
    k, e1, e1_fd, hgap, fint = (mp.mpf(v) for v in (k, e1, e1_fd, hgap, fint))
    corr = 2 * k * hgap * fint
    r21 = k * mp.tan(e1)
    e1v = e1 + e1_fd
    temp = corr / mp.cos(e1v) * (1 + mp.sin(e1v) ** 2)
    r43 = -k * mp.tan(e1v - temp)
    return r21, r43


EDGES = [
    dict(k=0.02, e1=0.05, e1_fd=0.0, hgap=0.03, fint=0.5),
    dict(k=-0.013, e1=-0.11, e1_fd=0.02, hgap=0.05, fint=0.7),
    dict(k=0.004, e1=0.0, e1_fd=0.0, hgap=0.02, fint=0.4),       # r21 = 0, r43 from fringe only
    dict(k=0.03, e1=0.2, e1_fd=-0.03, hgap=0.0, fint=0.0, delta_taper=2e-3),
]


@pytest.mark.parametrize('kw', EDGES)
@pytest.mark.parametrize('math', ['xm', 'libm'])
def test_dipole_edge_linear_closed_form(kw, math):
    rng = np.random.default_rng(11)
    n = 64
    x = rng.normal(0, 5e-3, n)
    y = rng.normal(0, 5e-3, n)
    px = rng.normal(0, 1e-4, n)
    py = rng.normal(0, 1e-4, n)
    chi_q = rng.choice([0.5, 1.0, 2.0], n)
    chi_m = rng.choice([1.0, 2.0, 4.0], n)
    p = xt.Particles(p0c=7e9, x=x, px=px, y=y, py=py, delta=rng.normal(0, 1e-3, n),
                     charge_ratio=chi_q, mass_ratio=chi_m)
    edge = xt.DipoleEdge(**kw)
    line = xt.Line(elements=[edge])
    line.build_tracker(backend=OracleBackend(math=math))
    line.track(p)
    q = p.sorted_by_id()
    r21, r43 = _edge_coeffs(kw['k'], kw['e1'], kw['e1_fd'], kw['hgap'], kw['fint'])
    taper = 1 + mp.mpf(kw.get('delta_taper', 0.0))
    chi = q['chi']
    for i in range(n):
        ex_px = mp.mpf(px[i]) + mp.mpf(chi[i]) * r21 * taper * mp.mpf(x[i])
        ex_py = mp.mpf(py[i]) + mp.mpf(chi[i]) * r43 * taper * mp.mpf(y[i])
        # the kick is a few roundings from the exact value: r21/r43 in binary64
        # (tan of a rounded angle) and two products and a sum
        for got, ex in ((q['px'][i], ex_px), (q['py'][i], ex_py)):
            assert abs(mp.mpf(got) - ex) <= 4e-16 * abs(ex) + 1e-22, (got, ex)
    np.testing.assert_array_equal(q['x'], x)
    np.testing.assert_array_equal(q['y'], y)
    # backtracking flips r21 and r43: the round trip restores px and py
    line.track(p, backtrack=True)
    b = p.sorted_by_id()
    assert np.max(np.abs(b['px'] - px)) <= 1e-19
    assert np.max(np.abs(b['py'] - py)) <= 1e-19


def test_dipole_edge_backtrack_is_the_negated_kick():
    kw = EDGES[1]
    rng = np.random.default_rng(4)
    n = 16
    x, y = rng.normal(0, 5e-3, n), rng.normal(0, 5e-3, n)
    p = xt.Particles(p0c=7e9, x=x, y=y, px=np.zeros(n), py=np.zeros(n))
    line = xt.Line(elements=[xt.DipoleEdge(**kw)])
    line.build_tracker(backend=OracleBackend(math='libm'))
    line.track(p, backtrack=True)
    q = p.sorted_by_id()
    r21, r43 = _edge_coeffs(kw['k'], kw['e1'], kw['e1_fd'], kw['hgap'], kw['fint'])
    for i in range(n):
        assert abs(mp.mpf(q['px'][i]) + r21 * mp.mpf(x[i])) <= 4e-16 * abs(r21 * x[i])
        assert abs(mp.mpf(q['py'][i]) + r43 * mp.mpf(y[i])) <= 4e-16 * abs(r43 * y[i])
