"""Shared helpers for the tests (lattices of every element type, particle
sets, exact comparison of particle records)."""
import numpy as np

import xtrack_amd as xt

COMPARED = ('x', 'px', 'y', 'py', 'zeta', 'delta', 'ptau', 'rpp', 'rvv', 's', 'ax', 'ay',
            'at_element', 'at_turn', 'state', 'particle_id')


def mixed_line(n_cells=3, aperture=0.02, with_cavity=True):
    """Small ring exercising every element type and magnet model on the
    path: expanded/exact drifts, thin multipoles, thick quadrupoles
    (mat-kick-mat), bends (rot-kick-rot Yoshida-4 plus the other curved
    models), sextupoles, octupoles, cavity, both aperture types."""
    els = []
    for ic in range(n_cells):
        s = 1 if ic % 2 == 0 else -1
        els += [
            xt.Quadrupole(length=1.2, k1=0.11),
            xt.LimitRect(min_x=-aperture, max_x=aperture, min_y=-aperture, max_y=aperture),
            xt.Drift(length=0.7),
            xt.Bend(length=3.0, angle=0.05),
            xt.LimitEllipse(a=aperture, b=aperture * 0.9),
            xt.Drift(length=0.5, model='exact'),
            xt.Sextupole(length=0.4, k2=0.3 * s, k2s=0.05),
            xt.Multipole(knl=[0.0, 0.001 * s, 0.02, 0.5 * s], ksl=[0, 0, 0.01]),
            xt.Bend(length=2.0, angle=0.03, k1=0.001, model='mat-kick-mat'),
            xt.Quadrupole(length=1.2, k1=-0.11, integrator='teapot',
                          num_multipole_kicks=3, knl=[0, 0, 0.01]),
            xt.Octupole(length=0.3, k3=2.0 * s),
            xt.Bend(length=2.0, angle=0.03, model='rot-kick-rot-high-order'),
            xt.Bend(length=2.0, angle=0.02, model='drift-kick-drift-exact', k2=0.1),
            xt.Bend(length=2.0, angle=0.02, model='rot-kick-rot-low-order',
                    edge_entry_angle=0.01, edge_exit_angle=0.015, edge_entry_hgap=0.02,
                    edge_entry_fint=0.5),
            xt.Sextupole(length=0.3, k2=0.2, integrator='yoshida4', num_multipole_kicks=9),
            xt.Marker(),
        ]
    if with_cavity:
        els.append(xt.Cavity(voltage=3e6, frequency=400e6, lag=175))
    return xt.Line(elements=els)


def segment_line(n_blocks=4, aperture=0.01):
    """Runs of cheap elements (drifts, both aperture types, thin
    multipoles, markers, a thin sextupole) that the line code tracks as
    segments, cut by thick quadrupoles; the kicks and the tight apertures
    lose particles at every position inside the runs."""
    els = []
    for ib in range(n_blocks):
        s = 1 if ib % 2 == 0 else -1
        for k in range(3):
            els += [
                xt.Drift(length=0.8),
                xt.LimitRect(min_x=-aperture, max_x=aperture, min_y=-aperture, max_y=aperture),
                xt.Multipole(knl=[0.0, 0.1 * s * (-1) ** k, 0.5], ksl=[0.0, 0.01]),
                xt.Drift(length=0.6),
                xt.LimitEllipse(a=aperture, b=aperture * 0.8),
                xt.Marker(),
                xt.Multipole(knl=[0.0, 0.0, 2.0 * s]),
                xt.Drift(length=0.4, model='exact'),
                xt.LimitRect(min_x=-aperture * 0.9, max_x=aperture, min_y=-aperture,
                             max_y=aperture * 0.9),
            ]
        els.append(xt.Quadrupole(length=1.0, k1=0.05 * s))
    return xt.Line(elements=els)


def random_particles(n, seed=0, scale=1e-3, p0c=7e9, spread=1.0):
    rng = np.random.default_rng(seed)
    return xt.Particles(p0c=p0c,
                        x=rng.normal(0, scale * spread, n), px=rng.normal(0, scale * 0.1 * spread, n),
                        y=rng.normal(0, scale * spread, n), py=rng.normal(0, scale * 0.1 * spread, n),
                        zeta=rng.normal(0, 0.05, n), delta=rng.normal(0, 1e-3, n))


def guard_edge_particles(n, seed=0, p0c=7e9):
    """Beam-like particles mixed with values on both sides of the ranges in
    which the kernel replaces divisions by reciprocal multiplications:
    px = +0, -0, subnormal and around 2^-900; pz driven below 2^-16 by a
    large |py| or by delta close to -1; very low energy (small rvv); |x|,
    |y| large enough for x*x to overflow; coordinates of +0 and -0."""
    rng = np.random.default_rng(seed)
    x = rng.normal(0, 1e-3, n)
    px = rng.normal(0, 1e-4, n)
    y = rng.normal(0, 1e-3, n)
    py = rng.normal(0, 1e-4, n)
    delta = rng.normal(0, 1e-3, n)
    k = n // 12
    px[0 * k:1 * k] = 0.0
    px[1 * k:2 * k] = -0.0
    px[2 * k:3 * k] = 5e-324 * rng.choice([-1, 1], k)
    px[3 * k:4 * k] = 2.0 ** -899 * rng.uniform(0.5, 2, k) * rng.choice([-1, 1], k)
    px[4 * k:5 * k] = 2.0 ** -901 * rng.uniform(0.5, 2, k)
    py[5 * k:6 * k] = (1 - 10.0 ** rng.uniform(-13, -8, k)) * rng.choice([-1, 1], k)
    px[5 * k:6 * k] = 0.0
    delta[6 * k:7 * k] = -1 + 10.0 ** rng.uniform(-7, -3, k)
    delta[7 * k:8 * k] = -1 + 10.0 ** rng.uniform(-12, -8, k)
    x[8 * k:9 * k] = rng.uniform(-0.9, 0.9, k)
    px[9 * k:10 * k] = rng.uniform(-0.3, 0.3, k)
    x[10 * k:11 * k] = 10.0 ** rng.uniform(150, 300, k) * rng.choice([-1, 1], k)
    y[10 * k:11 * k] = 10.0 ** rng.uniform(150, 300, k)
    zeta = rng.normal(0, 0.05, n)
    for arr in (x, y, px, py, zeta):
        arr[11 * k:12 * k] = rng.choice([0.0, -0.0], k)
    return xt.Particles(p0c=p0c, x=x, px=px, y=y, py=py, delta=delta, zeta=zeta)


def compare_exact(a, b, fields=COMPARED):
    """a, b: Particles.  Returns list of (field, n_mismatch, example) for
    fields whose values differ: floating-point values must have the same
    bits (the sign of zero included); NaNs match NaNs whatever their payload."""
    sa, sb = a.sorted_by_id(), b.sorted_by_id()
    bad = []
    for ff in fields:
        va, vb = sa[ff], sb[ff]
        if va.dtype.kind == 'f':
            eq = (((va == vb) & (np.signbit(va) == np.signbit(vb)))
                  | (np.isnan(va) & np.isnan(vb)))
        else:
            eq = va == vb
        if not eq.all():
            ii = int(np.flatnonzero(~eq)[0])
            bad.append((ff, int((~eq).sum()), (ii, va[ii], vb[ii])))
    return bad


def max_rel_err(a, b, fields=('x', 'px', 'y', 'py', 'zeta', 'delta')):
    sa, sb = a.sorted_by_id(), b.sorted_by_id()
    out = {}
    for ff in fields:
        va, vb = sa[ff], sb[ff]
        m = np.isfinite(va) & np.isfinite(vb)
        den = np.maximum(np.abs(vb[m]), 1e-300)
        out[ff] = float(np.max(np.abs(va[m] - vb[m]) / den)) if m.any() else 0.0
    return out
