"""Dynamic-aperture scans with plain tracking.

Initial conditions on a polar grid of normalised transverse amplitudes (in
beam sigmas, x_norm = r cos(theta), y_norm = r sin(theta), px_norm =
py_norm = 0) at one or more momentum deviations are tracked for a number of
turns; a particle's survival turn is its turn count at loss, or the number
of turns tracked.  The dynamic aperture at an angle is the largest grid
amplitude below which every particle at that angle survived.

Normalised coordinates are mapped to physical ones with the linear
normal form: for each momentum deviation, the 4d closed orbit at that
delta plus W of the 4d one-turn matrix applied to the normalised
coordinates scaled by the square roots of the geometric emittances.  The
reference's ``Line.build_particles`` (line.py:2276) hands this to
xpart.build_particles, which is not part of the reference tree; this
mapping is checked instead by the closed-form aperture of a linear ring
(tests/test_da.py).
"""
import numpy as np

from .particles import Particles
from .twiss import (_reference, compute_one_turn_matrix_finite_differences,
                    find_closed_orbit, get_linear_normal_form)


def build_normalized_particles(line, x_norm, px_norm, y_norm, py_norm, delta, nemitt_x,
                               nemitt_y, zeta=0.0, particle_ref=None):
    """Particles at the given normalised coordinates (beam sigmas) and
    momentum deviations; the line needs a tracker (it computes the closed
    orbit and one-turn matrix at each distinct delta)."""
    ref = _reference(line, particle_ref)
    x_norm, px_norm, y_norm, py_norm, delta, zeta = np.broadcast_arrays(
        *[np.atleast_1d(np.asarray(v, dtype=np.float64))
          for v in (x_norm, px_norm, y_norm, py_norm, delta, zeta)])
    bg = float(ref._np('beta0')[0] * ref._np('gamma0')[0])
    sx, sy = np.sqrt(nemitt_x / bg), np.sqrt(nemitt_y / bg)
    phys = np.zeros((len(x_norm), 4))
    for d in np.unique(delta):
        co = find_closed_orbit(line, ref, method='4d', delta0=float(d))
        R, _ = compute_one_turn_matrix_finite_differences(line, ref, co, kill_cavity=True)
        W = get_linear_normal_form(R, only_4d_block=True)[0]
        sel = delta == d
        norm = np.stack([x_norm[sel] * sx, px_norm[sel] * sx, y_norm[sel] * sy,
                         py_norm[sel] * sy])
        phys[sel] = (co[:4, None] + W[:4, :4] @ norm).T
    return Particles(q0=ref.q0, mass0=ref.mass0, p0c=float(ref._np('p0c')[0]),
                     x=phys[:, 0], px=phys[:, 1], y=phys[:, 2], py=phys[:, 3],
                     zeta=np.array(zeta), delta=np.array(delta))


def polar_grid(r_max, n_r, n_angles, deltas=(0.0,)):
    """Amplitudes r_max/n_r .. r_max (n_r steps) x angles (k + 1/2) pi/(2
    n_angles), k < n_angles, x each delta; flattened arrays (r, theta,
    delta)."""
    r = np.linspace(r_max / n_r, r_max, n_r)
    th = (np.arange(n_angles) + 0.5) * (np.pi / 2) / n_angles
    dd = np.asarray(deltas, dtype=np.float64)
    R, TH, D = np.meshgrid(r, th, dd, indexing='ij')
    return R.ravel(), TH.ravel(), D.ravel()


def survival_turns(particles, num_turns):
    """Per particle (in particle_id order): the turn at which it was lost,
    or num_turns if it is still alive."""
    s = particles.sorted_by_id()
    return np.where(s['state'] > 0, num_turns, s['at_turn'])


def dynamic_aperture(r, theta, delta, surv, num_turns):
    """{(theta, delta): DA} with DA the largest amplitude below which every
    particle at that angle and delta survived num_turns (r_max when none
    was lost: a lower bound)."""
    out = {}
    for th in np.unique(theta):
        for d in np.unique(delta):
            sel = (theta == th) & (delta == d)
            rr, ss = r[sel], surv[sel]
            order = np.argsort(rr)
            rr, ss = rr[order], ss[order]
            lost = np.flatnonzero(ss < num_turns)
            out[(float(th), float(d))] = float(rr[-1] if len(lost) == 0 else
                                               (rr[lost[0] - 1] if lost[0] > 0 else 0.0))
    return out


def da_scan(line, r_max, n_r, n_angles, num_turns, nemitt_x, nemitt_y, deltas=(0.0,),
            particle_ref=None, optics_line=None):
    """Plain dynamic-aperture scan: build the grid, track it num_turns turns
    on `line`'s tracker and return the grid, survival turns and per-angle
    apertures.  `optics_line` (default: `line`) computes the normal form."""
    r, th, d = polar_grid(r_max, n_r, n_angles, deltas)
    p = build_normalized_particles(optics_line or line, r * np.cos(th), 0.0, r * np.sin(th),
                                   0.0, d, nemitt_x, nemitt_y, particle_ref=particle_ref)
    if line.tracker is not None and getattr(line.tracker.backend, 'space', None) == 'device':
        p.move(_context=line.tracker.backend.device)
    line.track(p, num_turns=num_turns)
    surv = survival_turns(p, num_turns)
    return {'r': r, 'theta': th, 'delta': d, 'survival': surv,
            'da': dynamic_aperture(r, th, d, surv, num_turns)}
