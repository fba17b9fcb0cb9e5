"""Linear optics from the tracker itself: closed orbit, one-turn matrix by
central finite differences, uncoupled Twiss parameters, tunes,
chromaticity and matched Gaussian distributions.

The reference obtains these through ``Line.twiss`` (xtrack/twiss/twiss.py:33,
closed orbit in twiss/closed_orbit.py:24, one-turn matrix by finite
differences in line.py:3324, normal form in linear_normal_form.py:61) and
``xpart.build_particles``.  Here only what the synthetic workloads need is
provided: every quantity comes from tracking a handful of particles through
the same line with the same tracker used for production.
"""
import numpy as np

from .particles import Particles

COORDS = ('x', 'px', 'y', 'py', 'zeta', 'delta')


def _track_points(line, ref, pts, kill_cavity=False):
    """Track the 6D points (rows of pts) one turn; return final 6D points."""
    p = Particles(q0=ref.q0, mass0=ref.mass0, p0c=float(ref._np('p0c')[0]),
                  x=pts[:, 0], px=pts[:, 1], y=pts[:, 2], py=pts[:, 3],
                  zeta=pts[:, 4], delta=pts[:, 5])
    if line.tracker is None:
        line.build_tracker()
    fl = line.tracker.track_flags
    old_kill = fl.XS_FLAG_KILL_CAVITY_KICK
    old_ign = fl.XS_FLAG_IGNORE_LOCAL_APERTURE
    fl.XS_FLAG_KILL_CAVITY_KICK = kill_cavity
    fl.XS_FLAG_IGNORE_LOCAL_APERTURE = True
    try:
        line.track(p, num_turns=1)
    finally:
        fl.XS_FLAG_KILL_CAVITY_KICK = old_kill
        fl.XS_FLAG_IGNORE_LOCAL_APERTURE = old_ign
    s = p.sorted_by_id()
    if np.any(s['state'] <= 0):
        raise RuntimeError('probe particle lost during optics computation')
    return np.stack([s[c] for c in COORDS], axis=1)


def one_turn_matrix(line, ref, orbit=None, steps=(1e-8, 1e-9, 1e-8, 1e-9, 1e-6, 1e-7),
                    kill_cavity=False):
    """6x6 one-turn matrix around `orbit` by central differences."""
    orbit = np.zeros(6) if orbit is None else np.asarray(orbit, dtype=float)
    pts = np.tile(orbit, (13, 1))
    for j in range(6):
        pts[1 + 2 * j, j] += steps[j]
        pts[2 + 2 * j, j] -= steps[j]
    out = _track_points(line, ref, pts, kill_cavity=kill_cavity)
    M = np.zeros((6, 6))
    for j in range(6):
        M[:, j] = (out[1 + 2 * j] - out[2 + 2 * j]) / (2 * steps[j])
    return M, out[0]


def closed_orbit_4d(line, ref, delta, n_iter=6):
    """Transverse fixed point at fixed delta (cavity off), Newton iteration."""
    co = np.zeros(6)
    co[5] = delta
    for _ in range(n_iter):
        M, out = one_turn_matrix(line, ref, co, kill_cavity=True)
        res = out[:4] - co[:4]
        dx = np.linalg.solve(M[:4, :4] - np.eye(4), -res)
        co[:4] += dx
        if np.max(np.abs(res)) < 1e-15:
            break
    return co


def twiss_2d(m2):
    c = 0.5 * (m2[0, 0] + m2[1, 1])
    if abs(c) >= 1:
        raise ValueError('linearly unstable plane')
    s = np.sign(m2[0, 1]) * np.sqrt(1 - c * c)
    beta = m2[0, 1] / s
    alpha = (m2[0, 0] - m2[1, 1]) / (2 * s)
    mu = np.arctan2(s, c)
    if mu < 0:
        mu += 2 * np.pi
    return beta, alpha, mu / (2 * np.pi)


def twiss(line, ref, delta=0.0):
    """Uncoupled Twiss at the line start: betx, alfx, qx, bety, alfy, qy,
    dx, dpx (dispersion from the 6x6 matrix), and the 6x6 matrix."""
    co = closed_orbit_4d(line, ref, delta)
    M, _ = one_turn_matrix(line, ref, co, kill_cavity=True)
    bx, ax, qx = twiss_2d(M[0:2, 0:2])
    by, ay, qy = twiss_2d(M[2:4, 2:4])
    disp = np.linalg.solve(np.eye(4) - M[:4, :4], M[:4, 5])
    return dict(betx=bx, alfx=ax, qx=qx, bety=by, alfy=ay, qy=qy,
                dx=disp[0], dpx=disp[1], R=M, co=co)


def chromaticity(line, ref, ddelta=1e-5):
    tp = twiss(line, ref, ddelta)
    tm = twiss(line, ref, -ddelta)
    return ((tp['qx'] - tm['qx']) / (2 * ddelta), (tp['qy'] - tm['qy']) / (2 * ddelta))


def matched_gaussian(line, ref, n, nemitt_x, nemitt_y, sigma_z, sigma_delta, seed=0,
                     tw=None):
    """Gaussian beam matched to the uncoupled Twiss and dispersion at the
    line start; (zeta, delta) uncorrelated Gaussians; numpy
    default_rng(seed)."""
    if tw is None:
        tw = twiss(line, ref)
    beta0 = float(ref._np('beta0')[0])
    gamma0 = float(ref._np('gamma0')[0])
    ex = nemitt_x / (beta0 * gamma0)
    ey = nemitt_y / (beta0 * gamma0)
    rng = np.random.default_rng(seed)
    xn, pxn, yn, pyn = rng.standard_normal((4, n))
    zeta = sigma_z * rng.standard_normal(n)
    delta = sigma_delta * rng.standard_normal(n)
    x = np.sqrt(tw['betx'] * ex) * xn + tw['dx'] * delta
    px = np.sqrt(ex / tw['betx']) * (pxn - tw['alfx'] * xn) + tw['dpx'] * delta
    y = np.sqrt(tw['bety'] * ey) * yn
    py = np.sqrt(ey / tw['bety']) * (pyn - tw['alfy'] * yn)
    return Particles(q0=ref.q0, mass0=ref.mass0, p0c=float(ref._np('p0c')[0]),
                     x=x, px=px, y=y, py=py, zeta=zeta, delta=delta)
