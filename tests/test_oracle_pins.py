"""Pins of the CPU oracle (oracle/xt_oracle.c) against what does not depend
on it: the reference's own expected values for this path, closed forms,
exact geometry computed in high precision, brute force and invariants.
Each test names the reference test or source it mirrors.  The oracle runs
in both builds (platform libm = the reference's CPU arithmetic, and the
project math library = what the GPU evaluates)."""
import math
from fractions import Fraction

import mpmath as mp
import numpy as np
import pytest

import xtrack_amd as xt
from oracle_backend import OracleBackend, oracle_math

MATHS = ['libm', 'xm']


def _line(elements, math_='libm'):
    line = xt.Line(elements=elements)
    line.build_tracker(backend=OracleBackend(math=math_))
    return line


# ---------------------------------------------------------------------------
# drift (xtrack tests/test_elements.py:524-551)
@pytest.mark.parametrize('m', MATHS)
def test_drift_expanded_and_exact_closed_form(m):
    p0 = xt.Particles(p0c=1e9, px=0.3, py=-0.1, delta=0.02)
    d = 1 + 0.02
    for model, xprime in (('expanded', 0.3 / d), ('exact', 0.3 / math.sqrt(d * d - 0.09 - 0.01))):
        line = _line([xt.Drift(length=1.), xt.Drift(length=2.), xt.Drift(length=3.)], m)
        for ee in line.elements:
            ee.model = model
        line.build_tracker(backend=OracleBackend(math=m))
        p = p0.copy()
        line.track(p)
        assert abs(p.x[0] - xprime * 6) <= 1e-14 * 6


@pytest.mark.parametrize('m', MATHS)
def test_drift_zeta_from_path_length(m):
    """zeta = s - beta0 c t: dzeta = L - L_path * beta0/beta (mpmath)."""
    mp.mp.prec = 200
    p = xt.Particles(p0c=2e9, px=2e-3, py=-1e-3, delta=3e-3, zeta=0.1)
    line = _line([xt.Drift(length=5.0, model='exact')], m)
    q = p.copy()
    line.track(q)
    m0 = mp.mpf(xt.PROTON_MASS_EV)
    p0c = mp.mpf(2e9)
    delta = mp.mpf(float(p.delta[0]))
    pc = p0c * (1 + delta)
    beta = pc / mp.sqrt(pc ** 2 + m0 ** 2)
    beta0 = p0c / mp.sqrt(p0c ** 2 + m0 ** 2)
    px, py = mp.mpf(2e-3), mp.mpf(-1e-3)
    pz = mp.sqrt((1 + delta) ** 2 - px ** 2 - py ** 2)
    lpath = 5 * (1 + delta) / pz
    zeta = mp.mpf(0.1) + 5 - lpath * beta0 / beta
    assert abs(q.zeta[0] - float(zeta)) < 1e-14   # cancellation in 1 - rv0v*(1+d)/pz
    assert abs(q.x[0] - float(5 * px / pz)) < 1e-17


# ---------------------------------------------------------------------------
# cavity (xtrack tests/test_elements.py:822-847)
@pytest.mark.parametrize('m', MATHS)
def test_cavity_energy_gain(m):
    line = _line([xt.Cavity(frequency=0, phase=np.pi / 2, voltage=30)], m)
    part = xt.Particles(p0c=1e9, delta=[0, 1e-2], zeta=[0, 0.2])
    part0 = part.copy()
    line.track(part)
    assert np.allclose(part.energy, part0.energy + 30, atol=5e-7, rtol=0)
    Pc = np.sqrt(part.energy ** 2 - part.mass0 ** 2)
    delta = Pc / part.p0c - 1
    beta = Pc / part.energy
    assert np.allclose(part.delta, delta, atol=1e-14, rtol=0)
    assert np.allclose(part.rpp, 1 / (1 + delta), atol=1e-14, rtol=0)
    assert np.allclose(part.rvv, beta / part.beta0, atol=1e-14, rtol=0)
    assert np.allclose(part.zeta / part.beta0, part0.zeta / part0.beta0, atol=1e-14, rtol=0)
    assert np.allclose((part.ptau - part0.ptau) * part0.p0c, 30, atol=1e-9, rtol=0)


@pytest.mark.parametrize('m', MATHS)
def test_cavity_phase_and_frequency(m):
    """dE = q V sin(lag*pi/180 - 2 pi f zeta / (beta0 c)) (mpmath)."""
    mp.mp.prec = 120
    V, f, lag = 2e6, 200e6, 30.0
    line = _line([xt.Cavity(voltage=V, frequency=f, lag=lag)], m)
    z = np.array([-0.3, -0.01, 0.0, 0.05, 0.4])
    part = xt.Particles(p0c=26e9, zeta=z)
    part0 = part.copy()
    line.track(part)
    beta0 = mp.mpf(float(part.beta0[0]))
    for i, zz in enumerate(z):
        dE = V * mp.sin(mp.mpf(lag) * mp.pi / 180 - 2 * mp.pi * f * mp.mpf(zz) / (beta0 * 299792458))
        got = (part.ptau[i] - part0.ptau[i]) * 26e9
        assert abs(got - float(dE)) < 2e-6 * V


# ---------------------------------------------------------------------------
# thin multipole: complex kick (readable statement: ducktrack/elements.py:135-171)
@pytest.mark.parametrize('m', MATHS)
def test_thin_multipole_complex_kick(m):
    mp.mp.prec = 200
    knl = [1e-4, 0.05, -0.3, 2.0, 15.0]
    ksl = [-2e-4, 0.01, 0.4, -1.0, 5.0]
    rng = np.random.default_rng(3)
    x, y = rng.uniform(-0.02, 0.02, 50), rng.uniform(-0.02, 0.02, 50)
    part = xt.Particles(p0c=1e10, x=x, y=y, px=1e-4, py=-2e-4)
    line = _line([xt.Multipole(knl=knl, ksl=ksl)], m)
    line.track(part)
    for i in range(50):
        z = mp.mpc(x[i], y[i])
        s = sum((mp.mpf(knl[n]) + 1j * mp.mpf(ksl[n])) * z ** n / mp.factorial(n)
                for n in range(5))
        assert abs(part.px[i] - float(mp.mpf(1e-4) - s.real)) < 1e-17 + 1e-15 * abs(s.real)
        assert abs(part.py[i] - float(mp.mpf(-2e-4) + s.imag)) < 1e-17 + 1e-15 * abs(s.imag)


# ---------------------------------------------------------------------------
# thick quadrupole (mat-kick-mat): linear map of the quadrupole Hamiltonian
@pytest.mark.parametrize('m', MATHS)
def test_thick_quadrupole_transfer_matrix(m):
    mp.mp.prec = 200
    L, k1 = 3.1, 0.0085
    for delta in (0.0, 2e-3, -3e-3):
        part = xt.Particles(p0c=7e12, x=[1e-3, 0.0], px=[0.0, 1e-5], y=[0.0, 2e-3],
                            py=[3e-6, 0.0], delta=delta)
        line = _line([xt.Quadrupole(length=L, k1=k1)], m)
        line.track(part)
        K = mp.mpf(k1) / (1 + mp.mpf(delta))
        sk = mp.sqrt(K)
        cx, sx = mp.cos(sk * L), mp.sin(sk * L) / sk
        cy, sy = mp.cosh(sk * L), mp.sinh(sk * L) / sk
        d1 = 1 + mp.mpf(delta)
        for i, (x0, px0, y0, py0) in enumerate(((1e-3, 0.0, 0.0, 3e-6), (0.0, 1e-5, 2e-3, 0.0))):
            xp0, yp0 = mp.mpf(px0) / d1, mp.mpf(py0) / d1
            x1 = x0 * cx + xp0 * sx
            px1 = (-K * x0 * sx + xp0 * cx) * d1
            y1 = y0 * cy + yp0 * sy
            py1 = (K * y0 * sy + yp0 * cy) * d1
            for got, ref in ((part.x[i], x1), (part.px[i], px1), (part.y[i], y1),
                             (part.py[i], py1)):
                assert abs(got - float(ref)) < 1e-16 + 1e-13 * abs(float(ref))


# ---------------------------------------------------------------------------
# sector bend (rot-kick-rot, Yoshida-4) vs the exact circle in a uniform field
def _exact_bend(L, h, x0, px0, delta):
    """Horizontal motion of a particle of momentum (1+delta)p0 in the uniform
    field that bends the reference onto a circle of radius 1/h; x outward,
    reference counter-clockwise around the origin.  Returns x1, px1 and the
    path length, computed in mpmath."""
    mp.mp.prec = 200
    rho = 1 / mp.mpf(h)
    d1 = 1 + mp.mpf(delta)
    r0 = mp.matrix([rho + mp.mpf(x0), 0])
    pz0 = mp.sqrt(d1 ** 2 - mp.mpf(px0) ** 2)
    u0 = mp.matrix([mp.mpf(px0), pz0]) / d1          # (e_x, e_s) at phi=0 = (X, Y)
    rp = rho * d1                                      # radius of the particle's circle
    c = r0 + rp * mp.matrix([-u0[1], u0[0]])          # centre on the left of motion
    th = mp.mpf(h) * L
    ex = mp.matrix([mp.cos(th), mp.sin(th)])
    es = mp.matrix([-mp.sin(th), mp.cos(th)])
    # intersection of |t ex - c| = rp with t > 0 (closest to the reference)
    b = -2 * (ex[0] * c[0] + ex[1] * c[1])
    cc = c[0] ** 2 + c[1] ** 2 - rp ** 2
    disc = mp.sqrt(b * b - 4 * cc)
    ts = [(-b + disc) / 2, (-b - disc) / 2]
    t = min(ts, key=lambda v: abs(v - rho))
    r1 = t * ex
    rad = r1 - c
    u1 = mp.matrix([-rad[1], rad[0]]) / rp             # counter-clockwise tangent
    px1 = d1 * (u1[0] * ex[0] + u1[1] * ex[1])
    a0 = mp.atan2((r0 - c)[1], (r0 - c)[0])
    a1 = mp.atan2(rad[1], rad[0])
    return t - rho, px1, rp * (a1 - a0)


@pytest.mark.parametrize('m', MATHS)
@pytest.mark.parametrize('L,angle,nkicks', [(14.3, 2 * math.pi / 1104, 0), (2.0, 0.1, 21)])
def test_sector_bend_vs_exact_circle(m, L, angle, nkicks):
    h = angle / L
    cases = [(0.0, 0.0, 0.0), (1e-3, 0.0, 0.0), (0.0, 2e-4, 0.0), (-2e-3, 1e-4, 1e-3),
             (5e-4, -3e-4, -2e-3)]
    x0 = np.array([c[0] for c in cases])
    px0 = np.array([c[1] for c in cases])
    dl = np.array([c[2] for c in cases])
    part = xt.Particles(p0c=7e12, x=x0, px=px0, delta=dl)
    part0 = part.copy()
    line = _line([xt.Bend(length=L, angle=angle, num_multipole_kicks=nkicks)], m)
    line.reset_s_at_end_turn = False
    line.track(part)
    # tolerances = Yoshida-4 truncation error of these bends (measured:
    # 4e-14 / 4.6e-11 m for the arc bend, 1.3e-10 / 7.6e-9 m for the strong one)
    tol, tolz = (1e-12, 1e-10) if nkicks == 0 else (1e-9, 2e-8)
    for i in range(len(cases)):
        x1, px1, path = _exact_bend(L, h, *cases[i])
        assert abs(part.x[i] - float(x1)) < tol, (i, part.x[i], float(x1))
        assert abs(part.px[i] - float(px1)) < tol, (i, part.px[i], float(px1))
        dz = L - float(path) / part0.rvv[i]
        assert abs((part.zeta[i] - part0.zeta[i]) - dz) < tolz
        assert abs(part.s[i] - L) < 1e-14 * L   # sum of the eight Yoshida sub-lengths


def test_bend_integrator_is_fourth_order():
    """Halving the Yoshida-4 slice length divides the error by ~2^4
    (track_magnet.h:228-274, coefficients from MAD-NG)."""
    L, angle = 2.0, 0.2
    errs = []
    for nk in (7, 14, 28):
        part = xt.Particles(p0c=7e12, x=1e-3, px=2e-4, delta=1e-3)
        _line([xt.Bend(length=L, angle=angle, num_multipole_kicks=nk)]).track(part)
        x1, _, _ = _exact_bend(L, angle / L, 1e-3, 2e-4, 1e-3)
        errs.append(abs(part.x[0] - float(x1)))
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert 12 < r1 < 20 and 12 < r2 < 20, errs


# ---------------------------------------------------------------------------
# apertures (xtrack tests/test_apertures.py:46-75): brute-force grids
@pytest.mark.parametrize('m', MATHS)
def test_limit_rect_and_ellipse_truth_tables(m):
    XX, YY = np.meshgrid(np.linspace(-30e-3, 30e-3, 121), np.linspace(-30e-3, 30e-3, 121))
    x, y = XX.ravel(), YY.ravel()
    pr = xt.Particles(x=x, y=y)
    pe = pr.copy()
    _line([xt.LimitRect(min_x=-23e-3, max_x=23e-3, min_y=-18e-3, max_y=18e-3)], m).track(pr)
    a, b = 23e-3, 18e-3
    _line([xt.LimitEllipse(a=a, b=b)], m).track(pe)
    sr, se = pr.sorted_by_id(), pe.sorted_by_id()
    inside_r = (x >= -23e-3) & (x <= 23e-3) & (y >= -18e-3) & (y <= 18e-3)
    assert np.array_equal(sr['state'] > 0, inside_r)
    # exact rational evaluation away from the boundary
    A, B = Fraction(a) * Fraction(a), Fraction(b) * Fraction(b)
    for i in range(len(x)):
        v = Fraction(x[i]) ** 2 * B + Fraction(y[i]) ** 2 * A - A * B
        if abs(float(v)) < 1e-12 * float(A * B):
            continue
        assert (se['state'][i] > 0) == (v <= 0)
    assert set(np.unique(se['state'])) == {0, 1}


# global aperture (xtrack tests/test_apertures.py:16-43)
def test_global_aperture_after_static_thick_elements():
    line = _line([xt.Drift(length=2), xt.Quadrupole(length=2, k1=0),
                  xt.Multipole(length=2, isthick=True)])
    for start, expected in ((1, -1), (2, 1), (0, -1)):
        p = xt.Particles(p0c=7e12, px=0.6)
        line.track(p, ele_start=start, num_elements=1)
        assert p.state[0] == expected


# ---------------------------------------------------------------------------
# turn-by-turn monitor (xtrack tests/test_tracker.py:577-602)
@pytest.mark.parametrize('ele_start,ele_stop,expected_x', [
    (None, None, [0, 0.005, 0.010, 0.015, 0.020, 0.025]),
    (None, 3, [0, 0.005, 0.010, 0.015, 0.020, 0.023]),
    (2, None, [0, 0.003, 0.008, 0.013, 0.018, 0.023]),
    (2, 3, [0, 0.003, 0.008, 0.013, 0.018, 0.021]),
    (3, 2, [0, 0.002, 0.007, 0.012, 0.017, 0.022, 0.024]),
])
@pytest.mark.parametrize('with_progress', [False, 1, 2, 3])
def test_tbt_monitor_series(ele_start, ele_stop, expected_x, with_progress):
    line = _line([xt.Drift(length=1)] * 5)
    p = xt.Particles(px=0.001)
    line.track(p, num_turns=5, turn_by_turn_monitor=True, ele_start=ele_start,
               ele_stop=ele_stop, with_progress=with_progress)
    rec = line.record_last_track.x
    assert rec.shape == (1, len(expected_x) - 1)
    assert np.allclose(np.concatenate([rec[0], p.x]), expected_x, atol=1e-16, rtol=0)


def test_tracking_with_progress_s_accumulates():
    """xtrack tests/test_tracker.py:558-574 (non-collective)."""
    for turns in (1, 5, 7):
        for wp in (False, True, 1, 2, 3):
            line = _line([xt.Drift(length=2) for _ in range(5)])
            line.reset_s_at_end_turn = False
            p = xt.Particles(x=[1e-3, 2e-3, 3e-3], p0c=7e12)
            line.track(p, num_turns=turns, with_progress=wp)
            assert np.all(p.at_turn == turns)
            assert np.allclose(p.s, 10 * turns, rtol=0, atol=1e-14)


# partial tracking (xtrack tests/test_tracker.py:152-340)
def test_partial_tracking_matrix():
    n_elem = 9
    line = _line([xt.Drift(length=1.) for _ in range(n_elem)])
    p0 = xt.Particles(x=[1e-3, -2e-3, 5e-3], y=[2e-3, -4e-3, 3e-3], zeta=1e-2, p0c=7e12)

    def run(expected_turn, expected_el, **kw):
        p = p0.copy()
        if 'ele_start' in kw:
            p.at_element = kw['ele_start']
            p.s = float(kw['ele_start'])
        line.track(p, turn_by_turn_monitor=True, **kw)
        assert np.all(p.at_turn == expected_turn) and np.all(p.at_element == expected_el)
        assert np.all(p.s == expected_el)
        nmon = expected_turn if expected_el == 0 else expected_turn + 1
        assert line.record_last_track.x.shape == (3, nmon)

    for turns in (1, 2, 10):
        run(turns, 0, num_turns=turns)
        for start in range(n_elem):
            run(turns, 0, num_turns=turns, ele_start=start)
            for stop in range(start + 1, n_elem):
                run(turns - 1, stop, num_turns=turns, ele_start=start, ele_stop=stop)
            for stop in range(start + 1):
                run(turns, stop, num_turns=turns, ele_start=start, ele_stop=stop)
        for stop in range(1, n_elem):
            run(turns - 1, stop, num_turns=turns, ele_stop=stop)
    for start in range(n_elem):
        for shift in range(1, 3 * n_elem + 1):
            t = (start + shift) // n_elem
            run(t, start + shift - n_elem * t, ele_start=start, num_elements=shift)


# ---------------------------------------------------------------------------
# invariants: symplectic transverse one-turn map (cavity off)
@pytest.mark.parametrize('m', MATHS)
def test_transverse_map_is_symplectic(m):
    from xtrack_amd import optics
    from helpers import mixed_line
    line = mixed_line(n_cells=2)
    line.build_tracker(backend=OracleBackend(math=m))
    ref = xt.Particles(p0c=7e9)
    co = optics.closed_orbit_4d(line, ref, 1e-4)
    M, _ = optics.one_turn_matrix(line, ref, co, kill_cavity=True,
                                  steps=(1e-7, 1e-8, 1e-7, 1e-8, 1e-6, 1e-7))
    J = M[:4, :4]
    S = np.array([[0, 1, 0, 0], [-1, 0, 0, 0], [0, 0, 0, 1], [0, 0, -1, 0]], dtype=float)
    assert np.max(np.abs(J.T @ S @ J - S)) < 1e-6
    assert abs(np.linalg.det(J) - 1) < 1e-6


def test_sextupole_equals_drift_kick_drift():
    """Magnet identity (xtrack tests/test_magnet.py:111-156): a thick
    sextupole is drift(L/2) + thin kick k2*L + drift(L/2)."""
    p = xt.Particles(p0c=5e10, x=[1e-3, -2e-3], px=[1e-4, 0], y=[2e-3, 1e-3], py=[0, -3e-5],
                     delta=[0, 1e-3])
    pa, pb = p.copy(), p.copy()
    _line([xt.Sextupole(length=0.4, k2=3.0, k2s=-1.0)]).track(pa)
    _line([xt.Drift(length=0.2), xt.Multipole(knl=[0, 0, 1.2], ksl=[0, 0, -0.4]),
           xt.Drift(length=0.2)]).track(pb)
    for ff in ('x', 'px', 'y', 'py', 'zeta'):
        assert np.allclose(getattr(pa, ff), getattr(pb, ff), rtol=1e-15, atol=1e-18), ff


# ---------------------------------------------------------------------------
# libm vs project math library on the thin FODO ring: the effect of the
# elementary-function library on tracking (losses must not change)
def test_libm_vs_project_math_thin_fodo():
    from xtrack_amd import lattices, optics
    res = {}
    for m in MATHS:
        line = lattices.thin_fodo_ring()
        line.build_tracker(backend=OracleBackend(math=m, num_threads=4))
        ref = line.particle_ref
        tw = optics.twiss(line, ref)
        p = optics.matched_gaussian(line, ref, 2000, 2.5e-6, 2.5e-6, 0.2, 1e-3, seed=0, tw=tw)
        line.track(p, num_turns=100)
        res[m] = p.sorted_by_id()
    a, b = res['libm'], res['xm']
    for ff in ('state', 'at_turn', 'at_element'):
        assert np.array_equal(a[ff], b[ff])
    for ff in ('x', 'px', 'y', 'py', 'zeta', 'delta'):
        scale = np.max(np.abs(a[ff]))
        assert np.max(np.abs(a[ff] - b[ff])) <= 1e-10 * scale, ff


# ---------------------------------------------------------------------------
def test_math_library_ulp_bounds():
    mp.mp.prec = 200
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.uniform(-1, 1, 300), rng.uniform(-100, 100, 300),
                         10 ** rng.uniform(7, 300, 100) * rng.choice([-1, 1], 100)])
    bounds = {0: (mp.sin, 0.8), 1: (mp.cos, 0.8), 5: (mp.tan, 2.0)}
    for fn, (f, bound) in bounds.items():
        y = oracle_math(fn, xs)
        for v, r in zip(xs, y):
            ref = f(mp.mpf(v))
            assert abs(mp.mpf(r) - ref) <= bound * np.spacing(abs(float(ref))), (fn, v)
    xs = rng.uniform(-700, 700, 400)
    for fn, (f, bound) in {2: (mp.sinh, 1.5), 3: (mp.cosh, 1.5), 4: (mp.exp, 1.0)}.items():
        y = oracle_math(fn, xs)
        for v, r in zip(xs, y):
            ref = f(mp.mpf(v))
            assert abs(mp.mpf(r) - ref) <= bound * np.spacing(abs(float(ref))), (fn, v)


def test_one_over_factorial_table_is_correctly_rounded():
    """The 1/n! constants used by the multipole kick equal the correctly
    rounded values (headers/factorial.h)."""
    import re
    src = open(__file__.replace('tests/test_oracle_pins.py', 'oracle/xt_oracle.c')).read()
    blk = src[src.index('static const double t[20]'):]
    vals = re.findall(r'0x[0-9a-f.]+p[+-]\d+', blk[:blk.index('};')])
    assert len(vals) == 20
    for n, v in enumerate(vals):
        assert float.fromhex(v) == float(Fraction(1, math.factorial(n)))


# ---------------------------------------------------------------------------
# backtracking (xtrack tests/test_elements.py:429-474, tests/test_magnet.py:96-108,
# tests/test_tracker.py:528-553)
_BT_ELEMENTS = [
    ('drift', lambda: xt.Drift(length=1.0), 1e-14),
    ('drift_exact', lambda: xt.Drift(length=1.0, model='exact'), 1e-14),
    ('multipole', lambda: xt.Multipole(knl=[2, 3]), 1e-14),
    ('cavity', lambda: xt.Cavity(voltage=3.0), 1e-14),
    ('limitrect', lambda: xt.LimitRect(min_x=5), 1e-14),
    ('limitellipse', lambda: xt.LimitEllipse(a=10, b=10), 1e-14),
    ('quadrupole', lambda: xt.Quadrupole(length=1.0, k1=0.3), 5e-14),
    ('sextupole', lambda: xt.Sextupole(length=0.5, k2=0.4), 5e-14),
    ('octupole', lambda: xt.Octupole(length=0.5, k3=2.0), 5e-14),
    ('bend', lambda: xt.Bend(length=2.0, angle=0.05), 5e-14),
    ('bend_k1', lambda: xt.Bend(length=2.0, angle=0.05, k1=0.01, model='mat-kick-mat'), 5e-14),
    ('bend_edges', lambda: xt.Bend(length=2.0, angle=0.05, edge_entry_angle=0.01,
                                   edge_exit_angle=0.02), 5e-14),
]


@pytest.mark.parametrize('name,mk,tol', _BT_ELEMENTS, ids=[e[0] for e in _BT_ELEMENTS])
def test_backtrack_round_trip(name, mk, tol):
    """Forward then backward through one element returns the particle to
    its start (the reference's round-trip tests: 1e-14, magnets x 5e-14)."""
    line = xt.Line(elements=[mk()])
    line.build_tracker(backend=OracleBackend())
    p = xt.Particles(p0c=25.92e9, x=1e-3, px=1e-5, y=-2e-3, py=-1.5e-5, delta=1e-2,
                     zeta=1.0)
    p0 = p.copy()
    line.track(p)
    line.track(p, backtrack=True)
    a, b = p.to_numpy(), p0.to_numpy()
    for k in ('x', 'px', 'y', 'py', 'zeta', 'delta'):
        assert abs(a[k][0] - b[k][0]) <= tol * max(1.0, abs(b[k][0])), (name, k)
    assert abs(a['s'][0]) < 1e-7
    if name != 'limitrect':
        assert a['at_element'][0] == 0 and a['state'][0] == 1


def test_backtrack_element_by_element_matches_forward():
    """ONE_TURN_EBE records of a forward turn and of the backtracked turn
    coincide element by element (xtrack tests/test_tracker.py:528-553,
    1e-10); the particle returns to its start."""
    from helpers import mixed_line
    line = mixed_line(n_cells=2, aperture=0.5, with_cavity=False)
    line.build_tracker(backend=OracleBackend())
    p = xt.Particles(p0c=7e12, x=1e-4, px=1e-6, y=2e-4, py=3e-6, zeta=0.01, delta=1e-4)
    p0 = p.copy()
    line.track(p, turn_by_turn_monitor='ONE_TURN_EBE')
    fwd = line.record_last_track
    line.track(p, backtrack=True, turn_by_turn_monitor='ONE_TURN_EBE')
    bwd = line.record_last_track
    for k in ('x', 'px', 'y', 'py', 'zeta', 'delta'):
        assert np.allclose(getattr(fwd, k), getattr(bwd, k), rtol=0, atol=1e-10), k
    a, b = p.to_numpy(), p0.to_numpy()
    for k in ('x', 'px', 'y', 'py', 'zeta', 'delta'):
        assert abs(a[k][0] - b[k][0]) < 1e-13, k
    assert a['at_turn'][0] == 0 and a['at_element'][0] == 0


def test_backtrack_bookkeeping():
    """tracker.py:629-642, 699-731: at_turn decreases by one per backtracked
    turn, at_element is the number of elements at the turn start and s the
    line length, and at_element counts down; a partial backward range."""
    line = xt.Line(elements=[xt.Drift(length=1.0) for _ in range(6)])
    line.build_tracker(backend=OracleBackend())
    p = xt.Particles(p0c=7e12, x=[1e-3, 2e-3], px=[1e-5, -1e-5])
    p0 = p.copy()
    line.track(p, num_turns=3)
    h = p.to_numpy()
    assert list(h['at_turn']) == [3, 3] and list(h['at_element']) == [0, 0]
    line.track(p, num_turns=3, backtrack=True)
    h = p.to_numpy()
    assert list(h['at_turn']) == [0, 0] and list(h['at_element']) == [0, 0]
    assert np.allclose(h['x'], p0.to_numpy()['x'], rtol=0, atol=1e-16)
    q = p0.copy()
    line.track(q, ele_start=2, ele_stop=5, backtrack=True)
    h = q.to_numpy()
    # elements 4, 3, 2 backwards, starting from at_element 0
    assert list(h['at_element']) == [-3, -3] and list(h['at_turn']) == [0, 0]
    assert np.allclose(h['s'], -3.0)
