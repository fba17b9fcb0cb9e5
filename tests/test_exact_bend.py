"""Exact dipole maps of the 'bend-kick-bend' model (curved and straight
frames, track_magnet_drift.h:271-394).

The oracle's restatement is pinned by what the mathematics fixes: the
design particle of a bend with k0 = h stays on the axis with zeta
unchanged; the maps converge to the Yoshida-4 integrators (rot-kick-rot,
drift-kick-drift-exact) with many kicks, which a dropped term or a wrong
sign fails; the transverse map at fixed delta is symplectic.  The kernel's
element code and the generated line code, compiled for the host, must
then give the oracle's bits."""
import numpy as np
import pytest

import xtrack_amd as xt
from helpers import random_particles, compare_exact
from oracle_backend import OracleBackend, HostKernelBackend, GeneratedHostBackend


def _track(elements, p, backend, num_turns=1):
    line = xt.Line(elements=elements)
    line.build_tracker(backend=backend)
    q = p.copy()
    line.track(q, num_turns=num_turns)
    return q.sorted_by_id()


def _beam(n=200, seed=3, scale=2e-3):
    rng = np.random.default_rng(seed)
    return xt.Particles(p0c=7e9, x=rng.normal(0, scale, n), px=rng.normal(0, scale / 10, n),
                        y=rng.normal(0, scale, n), py=rng.normal(0, scale / 10, n),
                        zeta=rng.normal(0, 0.01, n), delta=rng.normal(0, 1e-3, n))


def test_design_orbit_stays_on_axis():
    p = xt.Particles(p0c=7e9, x=[0.0], px=[0.0], y=[0.0], py=[0.0], zeta=[0.0], delta=[0.0])
    for math in ('xm', 'libm'):
        r = _track([xt.Bend(length=3.0, angle=0.2, model='bend-kick-bend')], p,
                   OracleBackend(math=math))
        for ff in ('x', 'px', 'y', 'py'):
            assert abs(r[ff][0]) < 1e-15, (math, ff, r[ff][0])
        assert abs(r['zeta'][0]) < 1e-15


@pytest.mark.parametrize('k0_scale', [1.0, 1.3, -0.4])
def test_curved_map_converges_to_integrator(k0_scale):
    p = _beam()
    h = 0.1 / 3.0

    def bend(**kw):
        return xt.Bend(length=3.0, angle=0.1, k0=k0_scale * h, **kw)

    exact = _track([bend(model='bend-kick-bend')], p, OracleBackend(math='libm'))
    errs = []
    for n in (16, 32):
        ref = _track([bend(model='rot-kick-rot', integrator='yoshida4',
                           num_multipole_kicks=n)], p, OracleBackend(math='libm'))
        errs.append(max(np.max(np.abs(exact[ff] - ref[ff])) for ff in
                        ('x', 'px', 'y', 'py', 'zeta')))
    # the difference vanishes with the step (at better than second order);
    # a dropped term or a wrong sign leaves a difference that does not
    assert errs[1] < 1e-8 and errs[1] < errs[0] / 4, errs


def test_straight_map_closed_form_and_convergence():
    p = _beam(seed=5)
    k0 = 0.004
    el = [xt.Bend(length=1.5, angle=0.0, k0=k0, model='bend-kick-bend')]
    exact = _track(el, p, OracleBackend(math='libm'))
    p0 = p.sorted_by_id()
    # uniform field in a straight frame: px changes by exactly -k0 L
    assert np.allclose(exact['px'], p0['px'] - k0 * 1.5, rtol=0, atol=1e-17)
    ref = _track([xt.Bend(length=1.5, angle=0.0, k0=k0, model='drift-kick-drift-exact',
                          integrator='yoshida4', num_multipole_kicks=32)],
                 p, OracleBackend(math='libm'))
    for ff in ('x', 'px', 'y', 'py', 'zeta'):
        assert np.max(np.abs(exact[ff] - ref[ff])) < 1e-12, ff


@pytest.mark.parametrize('angle,k0', [(0.1, 'from_h'), (0.1, 0.05), (0.0, 0.004)])
def test_transverse_map_is_symplectic(angle, k0):
    el = [xt.Bend(length=2.0, angle=angle, k0=k0, model='bend-kick-bend')]
    z0 = np.array([1e-3, 2e-4, -5e-4, 1e-4])
    eps = 1e-5     # the straight map divides a difference of momenta by k0
    cols = []
    for sgn in (1, -1):
        xs = np.tile(z0, (4, 1)) + sgn * eps * np.eye(4)
        p = xt.Particles(p0c=7e9, x=xs[:, 0], px=xs[:, 1], y=xs[:, 2], py=xs[:, 3],
                         delta=np.full(4, 2e-4))
        r = _track(el, p, OracleBackend(math='libm'))
        cols.append(np.stack([r['x'], r['px'], r['y'], r['py']]))
    J = (cols[0] - cols[1]) / (2 * eps)
    S = np.kron(np.eye(2), np.array([[0.0, 1.0], [-1.0, 0.0]]))
    assert np.max(np.abs(J.T @ S @ J - S)) < 1e-8


def _bkb_ring():
    return xt.Line(elements=[
        xt.Quadrupole(length=1.2, k1=0.11),
        xt.LimitRect(min_x=-0.02, max_x=0.02, min_y=-0.02, max_y=0.02),
        xt.Drift(length=0.7),
        xt.Bend(length=3.0, angle=0.05, model='bend-kick-bend'),
        xt.Bend(length=2.0, angle=0.03, k1=0.002, k2=0.05, num_multipole_kicks=3,
                knl=[0, 0, 0.01], model='bend-kick-bend'),
        xt.Quadrupole(length=1.2, k1=-0.11),
        xt.Bend(length=1.5, angle=0.0, k0=0.004, model='bend-kick-bend'),
        xt.Bend(length=1.0, angle=0.002, k0=0.0, model='bend-kick-bend'),
        xt.Bend(length=1.0, angle=-0.002, k0=1e-12, model='bend-kick-bend'),
        xt.Sextupole(length=0.3, k2=0.2),
        xt.LimitEllipse(a=0.02, b=0.018),
        xt.Cavity(voltage=3e6, frequency=400e6, lag=175),
    ])


def _ions(n=400, seed=9):
    rng = np.random.default_rng(seed)
    return xt.Particles(p0c=7e9, x=rng.normal(0, 3e-3, n), px=rng.normal(0, 3e-4, n),
                        y=rng.normal(0, 3e-3, n), py=rng.normal(0, 3e-4, n),
                        zeta=rng.normal(0, 0.05, n), delta=rng.normal(0, 1e-3, n),
                        charge_ratio=rng.choice([0.5, 1.0, 2.0], n),
                        mass_ratio=rng.choice([1.0, 2.0, 4.0], n))


@pytest.mark.parametrize('backend', [HostKernelBackend, GeneratedHostBackend],
                         ids=['kernel_code', 'generated_code'])
def test_host_code_bit_exact(backend):
    for p in (random_particles(600, seed=2, scale=4e-3), _ions()):
        lh = _bkb_ring()
        lh.build_tracker(backend=backend())
        lc = _bkb_ring()
        lc.build_tracker(backend=OracleBackend())
        ph, pc = p.copy(), p.copy()
        lh.track(ph, num_turns=8)
        lc.track(pc, num_turns=8)
        assert not compare_exact(ph, pc)
        assert (pc.sorted_by_id()['state'] <= 0).any()


def test_xm_close_to_libm():
    p = random_particles(300, seed=4, scale=1e-3)
    a = _track(_bkb_ring().elements, p, OracleBackend(math='xm'), num_turns=1)
    b = _track(_bkb_ring().elements, p, OracleBackend(math='libm'), num_turns=1)
    alive = (a['state'] > 0) & (b['state'] > 0)
    assert np.array_equal(a['state'], b['state']) and alive.sum() > 100
    for ff in ('x', 'px', 'y', 'py', 'zeta'):
        assert np.max(np.abs(a[ff][alive] - b[ff][alive])) < 1e-12, ff


def test_atan2_asin_within_4_ulp():
    """The project's atan2 and asin (used by the exact maps) against the
    platform libm: at most 4 ulp apart, the same special values."""
    from oracle_backend import oracle_math
    rng = np.random.default_rng(17)
    n = 100000
    a = rng.normal(size=n) * 10 ** rng.uniform(-30, 30, n)
    b = rng.normal(size=n) * 10 ** rng.uniform(-30, 30, n)
    sp = np.array([0.0, -0.0, 1.0, -1.0, np.inf, -np.inf, np.nan, 1e-300])
    a[:64], b[:64] = np.repeat(sp, 8), np.tile(sp, 8)
    z = np.concatenate([rng.uniform(-1, 1, n), 1 - 10 ** rng.uniform(-16, -1, 1000),
                        [0.0, -0.0, 1.0, -1.0, 1.0 + 2e-16, np.nan]])
    for got, want in ((oracle_math(11, np.concatenate([a, b]), 'xm'), np.arctan2(a, b)),
                      (oracle_math(12, z, 'xm'), np.arcsin(z))):
        assert np.array_equal(np.isnan(got), np.isnan(want))
        ok = ~np.isnan(want)
        assert np.array_equal(np.signbit(got[ok]), np.signbit(want[ok]))
        ulp = np.abs(got[ok] - want[ok]) / np.spacing(np.abs(want[ok]))
        assert ulp.max() <= 4.0, ulp.max()


@pytest.mark.parametrize('kw', [dict(k1=0.002), dict(k1=-0.003, knl=[0, 0, 0.01]),
                                dict(k0=0.05 / 3.0 * 1.2, k1=0.001, k2=0.05)])
def test_curved_map_with_gradient_converges(kw):
    """Bend-kick-bend with a gradient (and multipoles): the exact dipole map
    between kicks of k1 (and knl) with k0_h_correction off.  With the kick
    count growing it converges to the Yoshida-4 rot-kick-rot integrator of
    the same magnet; a wrong sign or a dropped gradient term in the kick
    split does not converge."""
    p = _beam(seed=12)

    def bend(**extra):
        return xt.Bend(length=3.0, angle=0.05, **kw, **extra)

    exact = _track([bend(model='bend-kick-bend', num_multipole_kicks=32)], p,
                   OracleBackend(math='libm'))
    errs = []
    for n in (32, 64, 128):
        ref = _track([bend(model='rot-kick-rot', integrator='yoshida4',
                           num_multipole_kicks=n)], p, OracleBackend(math='libm'))
        errs.append(max(np.max(np.abs(exact[ff] - ref[ff])) for ff in
                        ('x', 'px', 'y', 'py', 'zeta')))
    # fourth-order convergence of the integrator onto the exact-map result
    # (factor 16 per halving of the step; >= 8 required)
    assert errs[2] < 1e-11 and errs[1] < errs[0] / 8 and errs[2] < errs[1] / 8, errs


@pytest.mark.parametrize('math', ['xm', 'libm'])
def test_exact_maps_backtrack_round_trip(math):
    """Forward through the bend-kick-bend ring (curved and straight exact
    maps, with and without gradient) and back: the initial coordinates
    within 5e-13 (track_magnet.h:414-442 reverses the kick sequence and
    negates the lengths)."""
    p = random_particles(300, seed=21, scale=2e-3)
    line = _bkb_ring()
    line.build_tracker(backend=OracleBackend(math=math))
    q = p.copy()
    line.track(q)
    line.track(q, backtrack=True)
    a, b = q.sorted_by_id(), p.sorted_by_id()
    alive = a['state'] > 0
    assert alive.sum() > 250
    for ff in ('x', 'px', 'y', 'py', 'zeta', 'delta'):
        assert np.max(np.abs(a[ff][alive] - b[ff][alive])) < 5e-13, ff


def test_exact_maps_backtrack_host_kernel_bit_exact():
    """The kernel's backtracking programs of the exact maps, compiled for the
    host, give the oracle's bits."""
    p = random_particles(400, seed=22, scale=3e-3)
    lh = _bkb_ring()
    lh.build_tracker(backend=HostKernelBackend())
    lc = _bkb_ring()
    lc.build_tracker(backend=OracleBackend())
    ph, pc = p.copy(), p.copy()
    for ll, pp in ((lh, ph), (lc, pc)):
        ll.track(pp, num_turns=3)
        ll.track(pp, num_turns=2, backtrack=True)
    assert not compare_exact(ph, pc)
