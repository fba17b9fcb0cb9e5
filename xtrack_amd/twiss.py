"""Twiss through the tracker: closed orbit, one-turn matrix, linear normal
form and the optics functions along the line.

Mirrors ``Line.twiss`` of the reference for the uncoupled-or-coupled linear
optics it derives from tracking (xtrack/twiss/twiss.py:33-760; closed
orbit twiss/closed_orbit.py:24-160; one-turn matrix by central finite
differences line.py:3324; normal form linear_normal_form.py:61-180).
Every number comes from tracking a few probe particles with the line's own
tracker (the HIP kernels on the GPU), so the optics are those of the maps
the production tracking uses.

Conventions: the canonical coordinates (x, px, y, py, zeta, pzeta) with
pzeta = ptau / beta0, as the reference's R_matrix and W_matrix use; the
symplectic form S pairs (x, px), (y, py), (zeta, pzeta).  The one-turn matrix is
R = W Rot W^-1 with Rot block-diagonal clockwise rotations by 2 pi q, and
W normalised so that W[0,1] = W[2,3] = W[4,5] = 0 (Courant-Snyder phase).
Along the line, W(s) = M(0->s) W(0) gives beta = W00^2 + W01^2,
alpha = -(W00 W10 + W01 W11), mu = atan2(W01, W00) / (2 pi).
"""
import numpy as np

from .particles import Particles

COORDS = ('x', 'px', 'y', 'py', 'zeta', 'pzeta')

S = np.array([[0., 1., 0., 0., 0., 0.],
              [-1., 0., 0., 0., 0., 0.],
              [0., 0., 0., 1., 0., 0.],
              [0., 0., -1., 0., 0., 0.],
              [0., 0., 0., 0., 0., 1.],
              [0., 0., 0., 0., -1., 0.]])

# finite-difference steps of the one-turn matrix (as the reference's
# defaults: dx 1e-7, dpx 1e-10, dy 1e-7, dpy 1e-10, dzeta 1e-6, ddelta 1e-7;
# the last is applied to pzeta)
STEPS_R_MATRIX = dict(dx=1e-7, dpx=1e-10, dy=1e-7, dpy=1e-10, dzeta=1e-6, ddelta=1e-7)


def _reference(line, particle_ref):
    ref = particle_ref if particle_ref is not None else line.particle_ref
    if ref is None:
        ref = getattr(line, 'json_particle', None)
    if ref is None:
        raise ValueError('a reference particle is needed (particle_ref=, or '
                         'line.particle_ref)')
    return ref


def _probes(ref, pts):
    """Particles at the canonical 6D points `pts` (rows) with the
    reference's p0c, mass0 and q0."""
    pts = np.atleast_2d(np.asarray(pts, dtype=np.float64))
    return Particles(q0=ref.q0, mass0=ref.mass0, p0c=float(ref._np('p0c')[0]),
                     x=pts[:, 0], px=pts[:, 1], y=pts[:, 2], py=pts[:, 3],
                     zeta=pts[:, 4], pzeta=pts[:, 5])


def pzeta_of_delta(ref, delta):
    """pzeta of a particle with the reference's energy and this delta."""
    p = Particles(q0=ref.q0, mass0=ref.mass0, p0c=float(ref._np('p0c')[0]),
                  delta=np.atleast_1d(np.asarray(delta, dtype=np.float64)))
    return p._np('ptau') / p._np('beta0')


def _canonical(rec):
    return np.stack([rec['x'], rec['px'], rec['y'], rec['py'], rec['zeta'],
                     rec['ptau'] / rec['beta0']], axis=-1)


def _track(line, ref, pts, kill_cavity, ebe=False):
    """One turn of the probe points (local apertures ignored, as the
    reference's twiss does); the final points, and with ebe the records at
    every element (shape n_points x (n_elements + 1) per coordinate)."""
    if line.tracker is None:
        line.build_tracker()
    fl = line.tracker.track_flags
    saved = (fl.XS_FLAG_KILL_CAVITY_KICK, fl.XS_FLAG_IGNORE_LOCAL_APERTURE)
    fl.XS_FLAG_KILL_CAVITY_KICK = kill_cavity
    fl.XS_FLAG_IGNORE_LOCAL_APERTURE = True
    p = _probes(ref, pts)
    try:
        line.track(p, num_turns=1, turn_by_turn_monitor='ONE_TURN_EBE' if ebe else None)
    finally:
        fl.XS_FLAG_KILL_CAVITY_KICK, fl.XS_FLAG_IGNORE_LOCAL_APERTURE = saved
    s = p.sorted_by_id()
    if np.any(s['state'] <= 0):
        raise RuntimeError('probe particle lost during the optics computation')
    out = _canonical(s)
    if not ebe:
        return out
    mon = line.record_last_track
    rec = {c: np.asarray(getattr(mon, c)).reshape(len(pts), -1)
           for c in ('x', 'px', 'y', 'py', 'zeta', 'ptau', 'beta0', 'delta', 's')}
    return out, rec


def compute_one_turn_matrix_finite_differences(line, particle_ref=None, particle_on_co=None,
                                               steps_r_matrix=None, kill_cavity=False):
    """6x6 one-turn matrix around the 6D point `particle_on_co` by central
    differences (line.py:3324); returns (R, the tracked centre point)."""
    ref = _reference(line, particle_ref)
    steps = dict(STEPS_R_MATRIX)
    steps.update(steps_r_matrix or {})
    h = np.array([steps['dx'], steps['dpx'], steps['dy'], steps['dpy'], steps['dzeta'],
                  steps['ddelta']])
    co = np.zeros(6) if particle_on_co is None else np.asarray(particle_on_co, dtype=float)
    pts = np.tile(co, (13, 1))
    for j in range(6):
        pts[1 + 2 * j, j] += h[j]
        pts[2 + 2 * j, j] -= h[j]
    out = _track(line, ref, pts, kill_cavity)
    R = np.zeros((6, 6))
    for j in range(6):
        R[:, j] = (out[1 + 2 * j] - out[2 + 2 * j]) / (2 * h[j])
    return R, out[0]


def find_closed_orbit(line, particle_ref=None, method='6d', delta0=0.0, co_guess=None,
                      tol=1e-14, max_iter=20, steps_r_matrix=None):
    """Fixed point of the one-turn map by Newton iteration with the
    finite-difference Jacobian (closed_orbit.py:24-160).  '4d': x, px, y,
    py at fixed delta = delta0 with the cavities off (zeta is not closed);
    '6d': all six coordinates with the cavities on.  Canonical coordinates
    (pzeta in the sixth place)."""
    ref = _reference(line, particle_ref)
    co = np.zeros(6) if co_guess is None else np.array(co_guess, dtype=float)
    if method == '4d':
        co[5] = pzeta_of_delta(ref, delta0)[0]
        idx, kill = slice(0, 4), True
    elif method == '6d':
        idx, kill = slice(0, 6), False
    else:
        raise ValueError(f'unknown method {method!r}')
    for _ in range(max_iter):
        R, out = compute_one_turn_matrix_finite_differences(
            line, ref, co, steps_r_matrix=steps_r_matrix, kill_cavity=kill)
        res = out[idx] - co[idx]
        n = 4 if method == '4d' else 6
        step = np.linalg.solve(R[:n, :n] - np.eye(n), -res)
        co[idx] += step
        if np.max(np.abs(res)) < tol or np.max(np.abs(step)) < tol:
            break
    return co


def _rot(mu):
    c, s = np.cos(mu), np.sin(mu)
    return np.array([[c, s], [-s, c]])


def get_linear_normal_form(R, only_4d_block=False):
    """R = W Rot W^-1 (linear_normal_form.py:61-180).  Returns W, invW,
    Rot and the three eigenvalues exp(2 pi i q) of the planes (x-like,
    y-like, zeta-like, assigned by the largest eigenvector components).
    only_4d_block: the longitudinal block is replaced by a rotation (fixed
    delta), as the reference does for the 4d method."""
    M = np.array(R, dtype=float)
    if only_4d_block:
        M[4:, :] = 0.0
        M[:, 4:] = 0.0
        M[4:, 4:] = _rot(np.pi / 10)
    w, v = np.linalg.eig(M)
    # keep the member of each conjugate pair with Im(v^H S v) > 0
    cands = []
    for k in range(6):
        vk = v[:, k]
        n = np.vdot(vk, S @ vk)      # purely imaginary for a symplectic M
        if n.imag > 0:
            cands.append((w[k], vk / np.sqrt(n.imag / 2.0)))
    if len(cands) != 3:
        raise ValueError('one-turn matrix is not stable (no three oscillation modes)')
    # assign modes to planes by the weight of their eigenvectors
    planes = [None, None, None]
    for lam, vk in sorted(cands, key=lambda c: -max(np.abs(c[1]))):
        weights = [np.linalg.norm(vk[2 * q:2 * q + 2]) for q in range(3)]
        for q in np.argsort(weights)[::-1]:
            if planes[q] is None:
                planes[q] = (lam, vk)
                break
    W = np.zeros((6, 6))
    eigs = np.zeros(3, dtype=complex)
    Rot = np.zeros((6, 6))
    for q, (lam, vk) in enumerate(planes):
        # phase such that the plane's position component is real positive
        vk = vk * np.exp(-1j * np.angle(vk[2 * q]))
        W[:, 2 * q] = vk.real
        W[:, 2 * q + 1] = vk.imag
        eigs[q] = lam
        Rot[2 * q:2 * q + 2, 2 * q:2 * q + 2] = _rot(np.angle(lam))
    invW = -S @ W.T @ S
    return W, invW, Rot, eigs


class TwissTable:
    """Columns (numpy arrays over the elements plus the end point, named
    like the reference's table) and scalars as attributes;
    ``tw['betx', 'ip1']`` selects by element name."""

    def __init__(self, columns, scalars):
        self._cols = columns
        self.__dict__.update(columns)
        self.__dict__.update(scalars)
        self._scalar_names = list(scalars)

    def __getitem__(self, key):
        if isinstance(key, tuple):
            col, name = key
            idx = list(self._cols['name']).index(name)
            return self._cols[col][idx]
        return self._cols[key]

    def keys(self):
        return list(self._cols) + self._scalar_names

    def cols(self):
        return list(self._cols)


def _tunes_at(line, ref, delta, steps_r_matrix):   # 4d tunes at this delta
    co = find_closed_orbit(line, ref, method='4d', delta0=delta,
                           steps_r_matrix=steps_r_matrix)
    R, _ = compute_one_turn_matrix_finite_differences(line, ref, co, steps_r_matrix,
                                                      kill_cavity=True)
    _, _, _, eigs = get_linear_normal_form(R, only_4d_block=True)
    return np.mod(np.angle(eigs[:2]) / (2 * np.pi), 1.0)


def twiss(line, particle_ref=None, method='6d', delta0=0.0, co_guess=None,
          steps_r_matrix=None, delta_disp=1e-6, eps_w=1e-7, compute_chromaticity=True,
          delta_chrom=1e-5):
    """Twiss table of a ring (twiss.py:33): closed orbit, one-turn matrix R,
    W, tunes (and qs for '6d'), then beta, alpha, gamma, phase advance,
    closed orbit and dispersion at every element, propagated by tracking
    the closed orbit with W-column and off-momentum probes through one
    turn with the element-by-element monitor."""
    ref = _reference(line, particle_ref)
    co = find_closed_orbit(line, ref, method=method, delta0=delta0, co_guess=co_guess,
                           steps_r_matrix=steps_r_matrix)
    kill = method == '4d'
    R, _ = compute_one_turn_matrix_finite_differences(line, ref, co, steps_r_matrix,
                                                      kill_cavity=kill)
    W, invW, Rot, eigs = get_linear_normal_form(R, only_4d_block=(method == '4d'))
    q = np.mod(np.angle(eigs) / (2 * np.pi), 1.0)

    # dispersion vector at the start: the off-momentum closed orbit slope
    # (cavities off), d co / d pzeta = (I - R4)^-1 R4[:4, 5]; the probes
    # step delta by +-delta_disp (pzeta by the corresponding amounts)
    if method == '4d':
        R4 = R
    else:
        R4, _ = compute_one_turn_matrix_finite_differences(line, ref, co, steps_r_matrix,
                                                           kill_cavity=True)
    d4 = np.linalg.solve(np.eye(4) - R4[:4, :4], R4[:4, 5])
    delta_co = float(_probes(ref, co)._np('delta')[0])
    pz_p, pz_m = pzeta_of_delta(ref, [delta_co + delta_disp, delta_co - delta_disp])

    # probes: co +- eps W[:, j] (transverse modes), off-momentum orbits
    pts = [co]
    for j in range(4):
        pts += [co + eps_w * W[:, j], co - eps_w * W[:, j]]
    for pz in (pz_p, pz_m):
        dp = pz - co[5]
        pts.append(co + dp * np.concatenate([d4, [0.0, 1.0]]))
    pts = np.array(pts)
    _, rec = _track(line, ref, pts, kill_cavity=True, ebe=True)
    can = {c: rec[c] for c in ('x', 'px', 'y', 'py', 'zeta')}
    can['pzeta'] = rec['ptau'] / rec['beta0']

    def col(c, k):
        return (can[c] if c in can else rec[c])[k]

    nrec = rec['x'].shape[1]
    Ws = np.zeros((nrec, 6, 4))
    for j in range(4):
        for ic, c in enumerate(COORDS):
            Ws[:, ic, j] = (col(c, 1 + 2 * j) - col(c, 2 + 2 * j)) / (2 * eps_w)
    out = {}
    names = list(line.element_names) + ['_end_point']
    out['name'] = np.array(names[:nrec], dtype=object)
    s = np.zeros(nrec)
    ss = 0.0
    for ii in range(nrec):
        s[ii] = ss
        if ii < len(line.elements) and line.elements[ii].isthick:
            ss += getattr(line.elements[ii], 'length', 0.0)
    out['s'] = s
    for c in ('x', 'px', 'y', 'py', 'zeta', 'pzeta', 'delta', 'ptau'):
        out[c] = col(c, 0)
    for pl, (i0, i1) in (('x', (0, 1)), ('y', (2, 3))):
        w00, w01 = Ws[:, i0, i0], Ws[:, i0, i1]
        w10, w11 = Ws[:, i1, i0], Ws[:, i1, i1]
        out['bet' + pl] = w00 ** 2 + w01 ** 2
        out['alf' + pl] = -(w00 * w10 + w01 * w11)
        out['gam' + pl] = w10 ** 2 + w11 ** 2
        mu = np.unwrap(np.arctan2(w01, w00))
        out['mu' + pl] = (mu - mu[0]) / (2 * np.pi)
    kp, km = 9, 10
    ddel = col('delta', kp) - col('delta', km)     # along the line, exact
    for c, dn in (('x', 'dx'), ('px', 'dpx'), ('y', 'dy'), ('py', 'dpy')):
        out[dn] = (col(c, kp) - col(c, km)) / ddel
    out['W_matrix'] = Ws
    scalars = dict(qx=q[0], qy=q[1], R_matrix=R, W_matrix0=W, particle_on_co=co,
                   circumference=line.get_length(), method=method, delta0=delta0)
    if method == '6d':
        # |arg| of the longitudinal eigenvalue: positive below and above
        # transition
        scalars['qs'] = abs(np.angle(eigs[2])) / (2 * np.pi)
    if compute_chromaticity:
        qp = _tunes_at(line, ref, delta0 + delta_chrom, steps_r_matrix)
        qm = _tunes_at(line, ref, delta0 - delta_chrom, steps_r_matrix)
        dq = (qp - qm + 0.5) % 1.0 - 0.5
        scalars['dqx'] = dq[0] / (2 * delta_chrom)
        scalars['dqy'] = dq[1] / (2 * delta_chrom)
    return TwissTable(out, scalars)
