"""Twiss through the tracker (xtrack_amd.twiss, the reference's
Line.twiss): closed orbit, one-turn matrix, normal form and optics
functions, pinned by thin-lens FODO closed forms, the closed-orbit response
to a dipole kick, symplecticity, the synchrotron tune of a thin cavity, and
the published optics of the reference's HL-LHC and SPS lattices."""
import math

import numpy as np
import pytest

import xtrack_amd as xt
from xtrack_amd import twiss as T
from oracle_backend import OracleBackend

L, F, N = 10.0, 8.0, 20     # half-cell drift, focal length, cells


def fodo_ring(kick=None):
    els, names = [], []
    for c in range(N):
        els += [xt.Multipole(knl=[0, 1 / F]), xt.Drift(length=L),
                xt.Multipole(knl=[0, -1 / F]), xt.Drift(length=L)]
        names += [f'qf{c}', f'da{c}', f'qd{c}', f'db{c}']
    if kick is not None:
        els.insert(2, xt.Multipole(knl=[kick]))
        names.insert(2, 'kicker')
    line = xt.Line(elements=els, element_names=names)
    line.build_tracker(backend=OracleBackend())
    return line


REF = xt.Particles(p0c=26e9)


def test_thin_fodo_closed_forms():
    """cos mu = 1 - L^2/(2 f^2); beta at the focusing / defocusing lens
    2L(1 +- sin(mu/2))/sin(mu); at a thin lens's entrance (where the table
    is taken) alpha = -+beta/(2f) (beta still rising into the focusing
    lens); Q = N mu / 2 pi."""
    tw = fodo_ring().twiss(REF, method='4d')
    mu = math.acos(1 - L * L / (2 * F * F))
    bmax = 2 * L * (1 + math.sin(mu / 2)) / math.sin(mu)
    bmin = 2 * L * (1 - math.sin(mu / 2)) / math.sin(mu)
    q = N * mu / (2 * math.pi)
    assert abs(tw.qx - q % 1) < 1e-12 and abs(tw.qy - q % 1) < 1e-12
    assert abs(tw.mux[-1] - q) < 1e-12 and abs(tw.muy[-1] - q) < 1e-12
    for c in (0, 7, 19):
        assert abs(tw['betx', f'qf{c}'] - bmax) < 1e-9 * bmax
        assert abs(tw['bety', f'qf{c}'] - bmin) < 1e-9 * bmax
        assert abs(tw['betx', f'qd{c}'] - bmin) < 1e-9 * bmax
        assert abs(tw['alfx', f'qf{c}'] + bmax / (2 * F)) < 1e-9 * bmax
        assert abs(tw['alfy', f'qf{c}'] - bmin / (2 * F)) < 1e-9 * bmax
    # gamma beta - alpha^2 = 1 everywhere
    assert np.allclose(tw.gamx * tw.betx - tw.alfx ** 2, 1.0, atol=1e-9)
    assert np.allclose(tw.s[-1], 2 * L * N)


def test_closed_orbit_response_to_a_dipole_kick():
    """x_co(s) = -theta sqrt(beta beta_k) cos(|mu - mu_k| - pi Q) / (2 sin pi Q)
    for a kick px -> px - theta at element k (linear optics)."""
    theta = 1e-7
    tw0 = fodo_ring(kick=0.0).twiss(REF, method='4d', compute_chromaticity=False)
    tw = fodo_ring(kick=theta).twiss(REF, method='4d', compute_chromaticity=False)
    k = list(tw0.name).index('kicker')
    Q = tw0.mux[-1]
    dmu = 2 * np.pi * np.abs(tw0.mux - tw0.mux[k])
    x_pred = -theta * np.sqrt(tw0.betx * tw0.betx[k]) * np.cos(dmu - np.pi * Q) / (
        2 * np.sin(np.pi * Q))
    assert np.max(np.abs(tw.x - x_pred)) < 1e-6 * np.max(np.abs(x_pred))
    assert np.max(np.abs(tw.x)) > 1e-7


def test_normal_form_and_one_turn_matrix_are_symplectic():
    from xtrack_amd import lattices
    line = lattices.collider_ring(n_arcs=1, n_cells=4, k2f=0.03, k2d=-0.06)
    line.build_tracker(backend=OracleBackend())
    tw = line.twiss(method='6d', compute_chromaticity=False)
    R, W = tw.R_matrix, tw.W_matrix0
    # finite differences with the reference's steps: entries up to ~3e2
    assert np.abs(R.T @ T.S @ R - T.S).max() < 1e-4
    # W is exactly normalised per mode; the cross terms carry the finite
    # differences' error
    assert np.abs(W.T @ T.S @ W - T.S).max() < 1e-7
    # R = W Rot W^-1
    _, invW, Rot, _ = T.get_linear_normal_form(R)
    # (to the finite differences' departure from symplecticity: the
    # eigenvalues sit off the unit circle by ~1e-7, Rot is exactly unitary)
    assert np.abs(W @ Rot @ invW - R).max() < 1e-6 * np.abs(R).max()


def test_synchrotron_tune_of_a_thin_cavity():
    """The ring starts with its cavity, so the one-turn matrix with the
    cavity on is the cavity-off matrix after the thin cavity's kick
    K: pzeta += a zeta, a = -(q V / p0c) (k / beta0^2) cos(lag)
    (track_rf.h:40-60 linearised).  The 6d normal form's qs must equal the
    longitudinal eigen-tune of R4 K."""
    from xtrack_amd import lattices
    line = lattices.collider_ring(n_arcs=1, n_cells=4, k2f=0.03, k2d=-0.06)
    line.build_tracker(backend=OracleBackend())
    ref = line.particle_ref
    assert isinstance(line.elements[0], xt.Cavity)
    tw4 = line.twiss(method='4d', compute_chromaticity=False)
    tw6 = line.twiss(method='6d', compute_chromaticity=False)
    cav = line.elements[0]
    beta0 = float(ref._np('beta0')[0])
    p0c = float(ref._np('p0c')[0])
    k = 2 * math.pi * cav.frequency / 299792458.0
    a = -(cav.voltage / p0c) * (k / beta0 ** 2) * math.cos(math.radians(cav.lag))
    K = np.eye(6)
    K[5, 4] = a
    ev = np.linalg.eigvals(tw4.R_matrix @ K)
    qs = np.min(np.abs(np.angle(ev))) / (2 * math.pi)
    assert 1e-3 < qs < 0.1
    assert abs(tw6.qs - qs) < 1e-6 * qs


@pytest.fixture(scope='module')
def hllhc_tw():
    from test_hllhc_lattice import hllhc_line, _load
    line = hllhc_line()
    line.build_tracker(backend=OracleBackend())
    pd = _load()['particle']
    return line.twiss(xt.Particles(p0c=pd['p0c'][0], mass0=pd['mass0']), method='4d')


def test_hllhc_optics(hllhc_tw):
    """HL-LHC v1.4 round collision optics: Q = 62.31 / 60.32, beta* = 15 cm
    at IP1 and IP5, +-250 urad half crossing angle (horizontal at IP1,
    vertical at IP5)."""
    tw = hllhc_tw
    assert abs(tw.mux[-1] - 62.31) < 2e-5 and abs(tw.muy[-1] - 60.32) < 2e-5
    for ip in ('ip1', 'ip5'):
        assert abs(tw['betx', ip] - 0.15) < 0.005
        assert abs(tw['bety', ip] - 0.15) < 0.005
    assert abs(tw['px', 'ip1'] - 250e-6) < 1e-7
    assert abs(tw['py', 'ip5'] - 250e-6) < 1e-7


def test_sps_tunes():
    from test_sps_lattice import sps_line
    line = sps_line()
    line.build_tracker(backend=OracleBackend())
    tw = line.twiss(line.json_particle, method='4d', compute_chromaticity=False)
    assert abs(tw.mux[-1] - 20.15) < 1e-6 and abs(tw.muy[-1] - 20.25) < 1e-6
