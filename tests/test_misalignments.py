"""Misaligned elements (track_misalignments.h, wrapped as in
track_local_particle_with_transformations.h:88-168): the oracle against the
geometry the transformations must reproduce, and the kernel's element code
(host builds) against the oracle bit for bit."""
import math

import numpy as np
import pytest

import xtrack_amd as xt
from helpers import compare_exact
from oracle_backend import OracleBackend, HostKernelBackend, GeneratedHostBackend

# misalignment of tests/test_misalign.py:90-101 in the reference
DX, DY, DS = -15.0, 7.0, 20.0
THETA, PHI, PSI, ANCHOR = -0.1, 0.11, 0.7, 8.0


def _line(elements, backend=None, global_aperture=True):
    line = xt.Line(elements=elements)
    line.build_tracker(backend=backend or OracleBackend())
    line.tracker.track_flags.XS_FLAG_IGNORE_GLOBAL_APERTURE = not global_aperture
    return line


def _track(elements, p, backend=None, global_aperture=True, **kw):
    line = _line(elements, backend, global_aperture)
    q = p.copy()
    line.track(q, **kw)
    return q.sorted_by_id()


def _p0():
    return xt.Particles(x=[0, 0.2, -0.4, -0.6, 0.8], y=[0, 0.2, 0.4, -0.6, -0.8],
                        px=[0, -0.01, -0.01, 0.01, 0.01], py=[0, -0.01, 0.01, -0.01, 0.01])


@pytest.mark.parametrize('angle,tilt', [(0, 0), (0.3, 0), (0, 0.1), (0.3, 0.1)],
                         ids=['straight', 'curved', 'straight-tilted', 'curved-tilted'])
def test_misaligned_field_free_element_is_invisible(angle, tilt):
    """A field-free element, however it is displaced and rotated, leaves
    the trajectory as the aligned one does: the entry transformation, the
    body in the misaligned frame and the exit transformation compose to the
    identity frame change (the reference's test_misalign_drift, tolerances
    atol 3e-14, rtol 1e-9).  The anchor ties the shifts to the rotation
    directions, so a sign error in any rotation breaks it."""
    length = 20.0
    if angle:
        mk = lambda **kw: xt.Bend(angle=angle, length=length, k0=0, rot_s_rad=tilt,
                                  model='rot-kick-rot', **kw)
    else:
        mk = lambda **kw: xt.Bend(angle=0.0, length=length, k0=0, rot_s_rad=tilt,
                                  model='drift-kick-drift-exact', **kw)
    # as in the reference test, the element alone: no global aperture (the
    # straight line leaves the curved frame by metres)
    p0 = _p0()
    ref = _track([mk()], p0, global_aperture=False)
    mis = dict(shift_x=DX, shift_y=DY, shift_s=DS, rot_y_rad=THETA, rot_x_rad=PHI,
               rot_s_rad_no_frame=PSI, rot_shift_anchor=ANCHOR)
    got = _track([mk(**mis)], p0, global_aperture=False)
    for ff in ('x', 'px', 'y', 'py', 's', 'zeta'):
        assert np.allclose(got[ff], ref[ff], atol=3e-14, rtol=1e-9), ff
    assert np.all(got['state'] == 1)
    # and backtracking through it restores the initial particles
    line = _line([mk(**mis)], global_aperture=False)
    q = p0.copy()
    line.track(q)
    line.track(q, backtrack=True)
    a, b = p0.sorted_by_id(), q.sorted_by_id()
    assert np.all(b['state'] == 1)
    for ff in ('x', 'px', 'y', 'py', 'zeta'):
        assert np.allclose(a[ff], b[ff], atol=1e-12, rtol=0), ff


def _exact_drift(c, length):
    """Drift_single_particle_exact, written out (textbook exact drift)."""
    opd = 1 + c['delta']
    pz = np.sqrt(opd * opd - c['px'] ** 2 - c['py'] ** 2)
    c['x'] = c['x'] + c['px'] / pz * length
    c['y'] = c['y'] + c['py'] / pz * length
    c['zeta'] = c['zeta'] + (1 - opd / (c['rvv'] * pz)) * length
    c['s'] = c['s'] + length


def _s_rotate(c, psi):
    cz, sz = math.cos(psi), math.sin(psi)
    x, y, px, py = c['x'], c['y'], c['px'], c['py']
    c['x'], c['y'] = cz * x + sz * y, -sz * x + cz * y
    c['px'], c['py'] = cz * px + sz * py, -sz * px + cz * py


@pytest.mark.parametrize('element_type', ['Quadrupole', 'Sextupole', 'Octupole', 'Multipole'])
def test_tilt_shifts_vs_sandwich(element_type):
    """shift_x/y/s and rot_s_rad of a straight element equal the sandwich
    exact drift(shift_s), translation, rotation, element, inverse rotation,
    inverse translation, exact drift(-shift_s) (the reference's
    test_tilt_shifts.py:17-103, same elements, tolerances 1e-13 and zeta
    5e-12)."""
    mk = {
        'Quadrupole': lambda **kw: xt.Quadrupole(k1=0.2, k1s=-0.3, length=3., **kw),
        'Sextupole': lambda **kw: xt.Sextupole(k2=0.1, k2s=0.2, length=0.3, **kw),
        'Octupole': lambda **kw: xt.Octupole(k3=0.1, k3s=0.2, length=0.4, **kw),
        'Multipole': lambda **kw: xt.Multipole(knl=[0.7, 0.8, 0.9, 1.0], ksl=[0.1, 0.2, 0.3, 0.4],
                                               length=0.4, hxl=0.1, **kw),
    }[element_type]
    sx, sy, ss, rs = 1e-3, 2e-3, 10e-3, -0.4
    p0 = xt.Particles(p0c=10e9, x=0.1, px=0.02, y=0.3, py=0.04, delta=0.03)
    flags = xt.TrackFlags()
    flags.XS_FLAG_IGNORE_GLOBAL_APERTURE = True

    got = _track([mk(shift_x=sx, shift_y=sy, shift_s=ss, rot_s_rad=rs)], p0)

    c = {k: np.array(v, dtype=np.float64) for k, v in p0.to_numpy().items()}
    _exact_drift(c, ss)
    c['x'], c['y'] = c['x'] - sx, c['y'] - sy
    _s_rotate(c, rs)
    mid = p0.copy()
    for ff in ('x', 'px', 'y', 'py', 'zeta', 's'):
        mid._np(ff)[:] = c[ff]
    mid = _track([mk()], mid)
    c = {k: np.array(mid[k], dtype=np.float64) for k in mid}
    _s_rotate(c, -rs)
    c['x'], c['y'] = c['x'] + sx, c['y'] + sy
    _exact_drift(c, -ss)
    for ff in ('x', 'px', 'y', 'py', 'delta'):
        assert np.allclose(got[ff], c[ff], rtol=0, atol=1e-13), ff
    assert np.allclose(got['zeta'], c['zeta'], rtol=0, atol=5e-12)

    # backtracking returns to the start (test_tilt_shifts.py:84-103)
    line = xt.Line(elements=[mk(shift_x=sx, shift_y=sy, shift_s=ss, rot_s_rad=rs)])
    line.build_tracker(backend=OracleBackend())
    q = p0.copy()
    line.track(q)
    line.track(q, backtrack=True)
    for ff in ('x', 'px', 'y', 'py', 'delta'):
        assert np.allclose(q._np(ff), p0._np(ff), rtol=0, atol=1e-11), ff
    assert np.allclose(q._np('zeta'), p0._np('zeta'), rtol=0, atol=1e-7)


def test_tilted_thin_dipole_kick():
    """A thin horizontal dipole kick rolled by psi with the frame kicks
    along (-cos psi, -sin psi) (rotation sign of test_rotation_signs.py)."""
    k0, psi = 1e-3, math.radians(30)
    p0 = xt.Particles(p0c=26e9, x=[0.0, 1e-3], y=[0.0, -2e-3])
    got = _track([xt.Multipole(knl=[k0], rot_s_rad=psi)], p0)
    assert np.allclose(got['px'], -k0 * math.cos(psi), rtol=1e-15, atol=0)
    assert np.allclose(got['py'], -k0 * math.sin(psi), rtol=1e-15, atol=0)
    assert np.allclose(got['x'], [0.0, 1e-3], rtol=0, atol=1e-19)
    assert np.allclose(got['y'], [0.0, -2e-3], rtol=0, atol=1e-19)


def test_rolled_bend():
    """rot_s_rad rolls the bend with its frame: a particle entering on the
    axis leaves it as through the unrolled bend, with the transverse plane
    turned by psi (the design orbit's small integrator offset turns with
    it); rot_s_rad_no_frame rolls the magnet only, and the field, now with
    a horizontal component, kicks the particle vertically by about
    -k0 L sin(psi)."""
    psi = 0.3
    p0 = xt.Particles(p0c=26e9, x=[0.0], y=[0.0], delta=[1e-3])
    flat = _track([xt.Bend(length=5.0, angle=0.02)], p0)
    got = _track([xt.Bend(length=5.0, angle=0.02, rot_s_rad=psi)], p0)
    c, s = math.cos(psi), math.sin(psi)
    assert abs(flat['x'][0]) > 1e-6 and flat['y'][0] == 0.0
    assert np.isclose(got['x'][0], c * flat['x'][0], rtol=1e-12, atol=1e-20)
    assert np.isclose(got['y'][0], s * flat['x'][0], rtol=1e-12, atol=1e-20)
    assert np.isclose(got['px'][0], c * flat['px'][0], rtol=1e-12, atol=1e-20)
    assert np.isclose(got['py'][0], s * flat['px'][0], rtol=1e-12, atol=1e-20)
    got = _track([xt.Bend(length=5.0, angle=0.02, rot_s_rad_no_frame=psi)], p0)
    assert np.isclose(got['py'][0], -0.02 * s, rtol=2e-2)


def _misaligned_line():
    return [
        xt.Drift(length=0.5),
        xt.Quadrupole(length=0.8, k1=0.4, shift_x=1e-3, shift_y=-2e-3, shift_s=0.01,
                      rot_s_rad=0.2, rot_x_rad=1e-3, rot_y_rad=-2e-3, rot_shift_anchor=0.3),
        xt.Sextupole(length=0.3, k2=2.0, shift_y=5e-4, rot_s_rad_no_frame=0.1),
        xt.Octupole(length=0.2, k3=50.0, shift_x=-3e-4),
        xt.Multipole(knl=[1e-4, 0.02], length=0.5, isthick=True, hxl=1e-4, shift_x=1e-3,
                     rot_y_rad=1e-3),
        xt.Bend(length=2.0, angle=0.02, rot_s_rad=0.5, shift_x=2e-4, rot_x_rad=5e-4,
                rot_shift_anchor=1.0),
        xt.Bend(length=1.0, angle=0.01, rot_s_rad=1.5707963267948966),
        xt.Cavity(voltage=5e6, frequency=400e6, lag=175, length=1.0, shift_x=1e-3,
                  rot_y_rad=1e-4),
        xt.DipoleEdge(k=0.01, e1=0.05, shift_y=1e-3, rot_s_rad=0.1),
        xt.LimitRect(min_x=-0.01, max_x=0.01, min_y=-0.01, max_y=0.01, shift_x=3e-3,
                     rot_s_rad=0.3),
        xt.LimitEllipse(a=0.01, b=0.008, shift_y=-2e-3, rot_x_rad=0.01),
        xt.Drift(length=0.5)]


@pytest.mark.parametrize('backend', [HostKernelBackend, GeneratedHostBackend],
                         ids=['interpreter', 'generated'])
def test_misaligned_elements_kernel_vs_oracle(backend):
    rng = np.random.default_rng(4)
    n = 600
    p = xt.Particles(p0c=7e9, x=rng.normal(0, 5e-3, n), px=rng.normal(0, 1e-4, n),
                     y=rng.normal(0, 5e-3, n), py=rng.normal(0, 1e-4, n),
                     zeta=rng.normal(0, 0.05, n), delta=rng.normal(0, 1e-3, n))
    # the library backtracks on the interpreter only
    for bt in ((False, True) if backend is HostKernelBackend else (False,)):
        lk = xt.Line(elements=_misaligned_line())
        lo = xt.Line(elements=_misaligned_line())
        lk.build_tracker(backend=backend())
        lo.build_tracker(backend=OracleBackend())
        pk, po = p.copy(), p.copy()
        lk.track(pk, num_turns=3, backtrack=bt)
        lo.track(po, num_turns=3, backtrack=bt)
        assert not compare_exact(pk, po)
        assert (po.state <= 0).sum() > 50
