"""RFMultipole (rf_multipole.py, track_rf.h:18-167): the oracle against the
closed form of an RF dipole kick and against symplecticity of the thin map
(the Panofsky-Wenzel relation between the transverse and the energy kick),
and the kernel's element code (host builds) against the oracle bit for
bit."""
import math

import numpy as np
import pytest

import xtrack_amd as xt
from helpers import compare_exact
from oracle_backend import OracleBackend, HostKernelBackend, GeneratedHostBackend

C_LIGHT = 299792458.0


def _track(elements, p, backend=None, **kw):
    line = xt.Line(elements=elements)
    line.build_tracker(backend=backend or OracleBackend())
    q = p.copy()
    line.track(q, **kw)
    return q.sorted_by_id()


def test_rf_dipole_closed_form():
    """Order-0 RF multipole without voltage: dpx = -k0 cos(phi),
    dptau = -(2 pi f / c) k0 x sin(phi), phi = phase_n - 2 pi f zeta/(beta0 c)."""
    k0, f, ph = 2e-4, 400.79e6, 0.3
    rng = np.random.default_rng(0)
    n = 50
    p = xt.Particles(p0c=7e12, x=rng.normal(0, 1e-3, n), y=rng.normal(0, 1e-3, n),
                     zeta=rng.normal(0, 0.1, n))
    el = xt.RFMultipole(frequency=f, knl=[k0], phase_n=[ph])
    got = _track([el], p)
    s = p.sorted_by_id()
    beta0 = s['beta0']
    phi = ph - 2 * math.pi * f * s['zeta'] / (beta0 * C_LIGHT)
    assert np.allclose(got['px'], -k0 * np.cos(phi), rtol=1e-13, atol=1e-22)
    assert np.allclose(got['py'], 0.0, atol=1e-22)
    dptau = -(2 * math.pi * f / C_LIGHT) * k0 * s['x'] * np.sin(phi)
    assert np.allclose(got['ptau'], dptau, rtol=1e-12, atol=1e-22)
    # the default order is 5 with zero-padded arrays
    assert xt.RFMultipole(knl=[1.0]).order == 5


def _map(el, z):
    """6D map in canonical (x, px, y, py, zeta, pzeta)."""
    p = xt.Particles(p0c=7e12, x=[z[0]], px=[z[1]], y=[z[2]], py=[z[3]], zeta=[z[4]],
                     pzeta=[z[5]])
    q = _track([el], p)
    return np.array([q['x'][0], q['px'][0], q['y'][0], q['py'][0], q['zeta'][0],
                     q['ptau'][0] / q['beta0'][0]])


def test_rf_dipole_symplectic():
    """The order-0 RF multipole with an accelerating voltage is a thin map
    derived from a potential, hence symplectic: J^T S J = S for the
    finite-difference Jacobian in (x, px, y, py, zeta, pzeta).  A sign or
    factor error between the transverse and the energy kick (the
    Panofsky-Wenzel relation) breaks it."""
    el = xt.RFMultipole(voltage=3e6, frequency=400e6, lag=20, knl=[1e-4], ksl=[2e-4],
                        phase_n=[0.1], phase_s=[0.5])
    z0 = np.array([2e-3, 1e-5, -1.5e-3, 2e-5, 0.05, 1e-4])
    h = np.array([1e-7, 1e-9, 1e-7, 1e-9, 1e-5, 1e-7])
    J = np.zeros((6, 6))
    for j in range(6):
        dz = np.zeros(6)
        dz[j] = h[j]
        J[:, j] = (_map(el, z0 + dz) - _map(el, z0 - dz)) / (2 * h[j])
    S = np.zeros((6, 6))
    for k in range(3):
        S[2 * k, 2 * k + 1], S[2 * k + 1, 2 * k] = 1.0, -1.0
    err = J.T @ S @ J - S
    assert np.abs(err).max() < 1e-9, np.abs(err).max()
    # the energy kick depends on x and y here
    assert abs(J[5, 0]) > 1e-6 and abs(J[5, 2]) > 1e-6


def test_rf_quadrupole_energy_kick_follows_tracking_code():
    """Order n >= 1: track_rf.h:96-118 multiplies knl[n]/n! by z^(n+1) in
    the energy kick, (n+1) times the term of the potential in the
    reference's own second implementation (ducktrack/elements.py:174-250),
    so the map is not symplectic for n >= 1.  The drop-in follows the
    tracking code: for an RF quadrupole dptau = -(2 pi f/c) k1 (x^2 - y^2)
    sin(phi), twice the potential's value."""
    k1, f, ph = 3e-3, 400e6, 0.4
    rng = np.random.default_rng(2)
    n = 30
    p = xt.Particles(p0c=7e12, x=rng.normal(0, 2e-3, n), y=rng.normal(0, 2e-3, n),
                     zeta=rng.normal(0, 0.1, n))
    got = _track([xt.RFMultipole(frequency=f, knl=[0, k1], phase_n=[0, ph])], p)
    s = p.sorted_by_id()
    phi = ph - 2 * math.pi * f * s['zeta'] / (s['beta0'] * C_LIGHT)
    x, y = s['x'], s['y']
    assert np.allclose(got['px'], -k1 * x * np.cos(phi), rtol=1e-12, atol=1e-22)
    assert np.allclose(got['py'], k1 * y * np.cos(phi), rtol=1e-12, atol=1e-22)
    dptau = -(2 * math.pi * f / C_LIGHT) * k1 * (x * x - y * y) * np.sin(phi)
    assert np.allclose(got['ptau'], dptau, rtol=1e-11, atol=1e-24)


def test_rf_phase_arguments_are_summed():
    """phase_n/phase_s (rad) add to pn/ps (deg), phase (rad) to lag (deg)
    (the reference's test_elements.py:121-215)."""
    knl, ksl = [0.1, 0.2], [-0.15, 0.05]
    pn, ps = np.array([10., 20.]), np.array([30., 40.])
    phase_n, phase_s = np.array([0.11, 0.22]), np.array([0.33, 0.44])
    p = xt.Particles(p0c=1e9, x=[1e-3], px=[2e-5], y=[-2e-3], py=[-3e-5], zeta=[0.02])
    a = _track([xt.RFMultipole(knl=knl, ksl=ksl, pn=pn, ps=ps, phase_n=phase_n,
                               phase_s=phase_s, frequency=400e6)], p)
    b = _track([xt.RFMultipole(knl=knl, ksl=ksl, pn=pn + np.rad2deg(phase_n),
                               ps=ps + np.rad2deg(phase_s), frequency=400e6)], p)
    for nn in ('x', 'px', 'y', 'py', 'zeta', 'delta'):
        assert np.allclose(a[nn], b[nn], rtol=1e-14, atol=1e-14), nn
    p = xt.Particles(p0c=1e9, zeta=[0.02])
    a = _track([xt.RFMultipole(knl=[0], voltage=1000, frequency=400e6, phase=0.23)], p)
    b = _track([xt.RFMultipole(knl=[0], voltage=1000, frequency=400e6,
                               lag=np.rad2deg(0.23))], p)
    assert np.allclose(a['delta'], b['delta'], rtol=1e-14, atol=1e-14)
    assert a['delta'][0] != 0.0


def test_rf_multipole_backtrack_round_trip():
    el = xt.RFMultipole(voltage=2e6, frequency=400e6, lag=10, knl=[1e-4, 2e-3, 0.1],
                        ksl=[0, 1e-3], phase_n=[0.1, 0.2], rot_s_rad=0.2, shift_x=1e-3)
    rng = np.random.default_rng(1)
    n = 40
    p = xt.Particles(p0c=7e12, x=rng.normal(0, 1e-3, n), y=rng.normal(0, 1e-3, n),
                     px=rng.normal(0, 1e-5, n), zeta=rng.normal(0, 0.1, n),
                     delta=rng.normal(0, 1e-4, n))
    line = xt.Line(elements=[el, xt.Drift(length=2.0)])
    line.build_tracker(backend=OracleBackend())
    q = p.copy()
    line.track(q)
    line.track(q, backtrack=True)
    a, b = p.sorted_by_id(), q.sorted_by_id()
    for ff in ('x', 'px', 'y', 'py', 'zeta', 'delta', 'ptau'):
        assert np.allclose(a[ff], b[ff], rtol=0, atol=1e-15), ff


def rf_line():
    return [
        xt.Drift(length=1.0),
        xt.RFMultipole(voltage=1e6, frequency=400e6, lag=30, knl=[1e-4, 2e-3, 0.1],
                       ksl=[0, 1e-3], phase_n=[0.1, 0.2], pn=[10.0]),
        xt.Quadrupole(length=1.0, k1=0.3),
        xt.RFMultipole(frequency=200e6, knl=[0, 0, 0, 5.0], ps=[45.0], ksl=[1e-4],
                       rot_s_rad=0.3, shift_y=1e-3),
        xt.LimitEllipse(a=0.01, b=0.01),
        xt.Drift(length=1.0)]


@pytest.mark.parametrize('backend', [HostKernelBackend, GeneratedHostBackend],
                         ids=['interpreter', 'generated'])
def test_rf_multipole_kernel_vs_oracle(backend):
    rng = np.random.default_rng(5)
    n = 400
    p = xt.Particles(p0c=7e9, x=rng.normal(0, 4e-3, n), px=rng.normal(0, 1e-4, n),
                     y=rng.normal(0, 4e-3, n), py=rng.normal(0, 1e-4, n),
                     zeta=rng.normal(0, 0.2, n), delta=rng.normal(0, 1e-3, n),
                     charge_ratio=rng.choice([1.0, 0.5], n),
                     mass_ratio=rng.choice([1.0, 2.0], n))
    for bt in ((False, True) if backend is HostKernelBackend else (False,)):
        lk, lo = xt.Line(elements=rf_line()), xt.Line(elements=rf_line())
        lk.build_tracker(backend=backend())
        lo.build_tracker(backend=OracleBackend())
        pk, po = p.copy(), p.copy()
        lk.track(pk, num_turns=3, backtrack=bt)
        lo.track(po, num_turns=3, backtrack=bt)
        assert not compare_exact(pk, po)
