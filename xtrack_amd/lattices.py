"""Synthetic lattices and beams of the benchmark workloads.

thin_fodo_ring: 100 cells of Drift(1 m) + thin quadrupole Multipole
    (knl[1] = +/-0.05 alternating) + LimitEllipse(a = b = 0.02 m), one
    Cavity (2 MV, 200 MHz, lag 180 deg); protons at p0c = 26 GeV.
collider_ring: 8 arcs x 23 FODO cells; a cell holds 6 thick Bends
    (14.3 m, angle 2 pi/1104), thick Quadrupoles (3.1 m, k1 = +/-0.0085),
    two Sextupoles (0.37 m) powered for chromaticity +2 in both planes and a
    thin octupole Multipole (knl[3] = +/-0.5); the 1.3 m gaps between
    magnets are drifts split in 6 pieces; a LimitRect (+/-0.022 x
    +/-0.018 m) follows every element; one Cavity (400 MHz, 16 MV, lag 180)
    at the start.  28,338 elements; protons at 7 TeV.
collider_with_errors: the collider ring with a seeded set of b3..b6
    multipole errors on every magnet (dynamic-aperture studies).
"""
import math

import numpy as np

from . import elements as xe
from .line import Line
from .particles import Particles
from .constants import PROTON_MASS_EV


def thin_fodo_ring():
    els, names = [], []
    els.append(xe.Cavity(voltage=2e6, frequency=200e6, lag=180.0))
    names.append('cav')
    for ic in range(100):
        k1l = 0.05 if ic % 2 == 0 else -0.05
        els += [xe.Drift(length=1.0), xe.Multipole(knl=[0.0, k1l]),
                xe.LimitEllipse(a=0.02, b=0.02)]
        names += [f'd{ic}', f'q{ic}', f'ap{ic}']
    line = Line(elements=els, element_names=names)
    line.particle_ref = Particles(p0c=26e9, mass0=PROTON_MASS_EV)
    return line


def collider_ring(n_arcs=8, n_cells=23, k2f=0.0, k2d=0.0, gap=1.3, n_gap_pieces=6,
                  bend_length=14.3, quad_length=3.1, k1=0.0085, sext_length=0.37,
                  oct_knl=0.5, aperture=(0.022, 0.018), cav_voltage=16e6,
                  cav_frequency=400e6):
    ax, ay = aperture
    n_bends = 6 * n_arcs * n_cells
    angle = 2 * math.pi / n_bends

    def ap():
        return xe.LimitRect(min_x=-ax, max_x=ax, min_y=-ay, max_y=ay)

    els, names = [xe.Cavity(voltage=cav_voltage, frequency=cav_frequency, lag=180.0), ap()], \
        ['cav', 'cav_ap']
    dl = gap / n_gap_pieces
    for ia in range(n_arcs):
        for ic in range(n_cells):
            tag = f'a{ia}c{ic}'
            sgn = 1.0 if (ia * n_cells + ic) % 2 == 0 else -1.0
            mags = [('qf', xe.Quadrupole(length=quad_length, k1=k1)),
                    ('sf', xe.Sextupole(length=sext_length, k2=k2f)),
                    ('b1', xe.Bend(length=bend_length, angle=angle)),
                    ('b2', xe.Bend(length=bend_length, angle=angle)),
                    ('b3', xe.Bend(length=bend_length, angle=angle)),
                    ('oc', xe.Multipole(knl=[0.0, 0.0, 0.0, sgn * oct_knl])),
                    ('qd', xe.Quadrupole(length=quad_length, k1=-k1)),
                    ('sd', xe.Sextupole(length=sext_length, k2=k2d)),
                    ('b4', xe.Bend(length=bend_length, angle=angle)),
                    ('b5', xe.Bend(length=bend_length, angle=angle)),
                    ('b6', xe.Bend(length=bend_length, angle=angle))]
            for mn, mm in mags:
                els += [mm, ap()]
                names += [f'{mn}.{tag}', f'{mn}.{tag}.ap']
                for ig in range(n_gap_pieces):
                    els += [xe.Drift(length=dl), ap()]
                    names += [f'd{ig}.{mn}.{tag}', f'd{ig}.{mn}.{tag}.ap']
    line = Line(elements=els, element_names=names)
    line.particle_ref = Particles(p0c=7e12, mass0=PROTON_MASS_EV)
    return line


def sextupole_fodo_ring(k2f=8.0, k2d=-5.0, every=10):
    """The thin FODO ring with a thin sextupole after every `every`-th
    quadrupole (alternating k2f, k2d): a fast ring with a dynamic aperture
    of a few sigma for scans and surrogate tests."""
    line = thin_fodo_ring()
    els, names = list(line.elements), list(line.element_names)
    for k in range(0, 100, every):
        i = names.index(f'q{k}')
        els.insert(i + 1, xe.Multipole(knl=[0.0, 0.0, k2f if (k // every) % 2 else k2d]))
        names.insert(i + 1, f'sx{k}')
    out = Line(elements=els, element_names=names)
    out.particle_ref = line.particle_ref
    return out


def collider_variant(seed, k2f=0.0, k2d=0.0, n_arcs=8, cells_per_arc=(8, 40), rel_err=1e-4,
                     **kwargs):
    """Lattice variant of the collider ring for batched studies: the number
    of cells per arc drawn uniformly in `cells_per_arc` (inclusive) and every
    quadrupole k1 and sextupole k2 multiplied by 1 + N(0, rel_err), all from
    numpy default_rng(seed)."""
    rng = np.random.default_rng(seed)
    n_cells = int(rng.integers(cells_per_arc[0], cells_per_arc[1] + 1))
    line = collider_ring(n_arcs=n_arcs, n_cells=n_cells, k2f=k2f, k2d=k2d, **kwargs)
    for ee in line.elements:
        if isinstance(ee, xe.Quadrupole):
            ee.k1 = ee.k1 * (1.0 + rel_err * rng.standard_normal())
        elif isinstance(ee, xe.Sextupole):
            ee.k2 = ee.k2 * (1.0 + rel_err * rng.standard_normal())
    return line


def collider_with_errors(seed, k2f=0.0, k2d=0.0, rel=1e-4, r0=0.017, orders=(3, 4, 5, 6),
                         **kwargs):
    """The collider ring with a random multipole error set: on every bend,
    quadrupole and sextupole, each normal multipole b_n of `orders`
    (European numbering: b3 sextupole ... b6 dodecapole) other than the
    magnet's main order gets a field error of N(0, rel) relative to the main
    field at the reference radius r0, i.e.

        knl[n-1] += eps * F_main(r0) * (n-1)! / r0**(n-1),

    with F_main(r0) = angle (bend), k1 L r0 (quadrupole), k2 L r0**2 / 2
    (sextupole) and eps ~ N(0, rel) from numpy default_rng(seed), drawn
    magnet by magnet in line order, orders ascending."""
    rng = np.random.default_rng(seed)
    line = collider_ring(k2f=k2f, k2d=k2d, **kwargs)
    for ee in line.elements:
        if isinstance(ee, xe.Bend):
            main, f_main = 1, ee.angle
        elif isinstance(ee, xe.Quadrupole):
            main, f_main = 2, ee.k1 * ee.length * r0
        elif isinstance(ee, xe.Sextupole):
            main, f_main = 3, ee.k2 * ee.length * r0 ** 2 / 2
        else:
            continue
        knl = np.array(ee.knl, dtype=np.float64)
        for n in orders:
            if n == main:
                continue
            knl[n - 1] += rel * rng.standard_normal() * f_main * math.factorial(n - 1) / \
                r0 ** (n - 1)
        ee.knl = knl
    return line


def set_sextupoles(line, k2f, k2d):
    for nn, ee in zip(line.element_names, line.elements):
        if nn.startswith('sf.'):
            ee.k2 = k2f
        elif nn.startswith('sd.'):
            ee.k2 = k2d


def match_chromaticity(make_line, target=(2.0, 2.0), backend_factory=None, n_iter=2):
    """Sextupole strengths (k2f, k2d) giving chromaticity `target`; the
    chromaticity is affine in (k2f, k2d) to first order, refined by Newton
    steps.  make_line(k2f, k2d) must return a line with particle_ref."""
    from .optics import chromaticity

    def chroma(k2f, k2d):
        line = make_line(k2f, k2d)
        line.build_tracker(backend=backend_factory() if backend_factory else None)
        return np.array(chromaticity(line, line.particle_ref))

    k = np.zeros(2)
    c0 = chroma(*k)
    for _ in range(n_iter):
        h = 0.01
        J = np.zeros((2, 2))
        J[:, 0] = (chroma(k[0] + h, k[1]) - c0) / h
        J[:, 1] = (chroma(k[0], k[1] + h) - c0) / h
        k = k + np.linalg.solve(J, np.asarray(target) - c0)
        c0 = chroma(*k)
    return k, c0
