"""Dynamic-aperture scan (xtrack_amd/da.py) on the CPU oracle: the
normalised-coordinate mapping and the aperture extraction against the
closed form of a linear ring, where a particle of normalised amplitude r at
angle theta survives exactly when r^2 eps (cos^2 theta beta_x + sin^2
theta beta_y) <= a^2 at every circular aperture of radius a (the betatron
phases of x and y reach their maxima together over the turns)."""
import numpy as np
import pytest

from xtrack_amd import lattices, da
from oracle_backend import OracleBackend


def test_linear_ring_closed_form():
    line = lattices.thin_fodo_ring()
    line.build_tracker(backend=OracleBackend())
    tw = line.twiss(method='4d')
    iap = [i for i, nn in enumerate(line.element_names) if nn.startswith('ap')]
    bx, by = np.asarray(tw['betx'])[iap], np.asarray(tw['bety'])[iap]
    ref = line.particle_ref
    geps = 2.5e-6 / float(ref._np('beta0')[0] * ref._np('gamma0')[0])
    step = 0.25
    res = da.da_scan(line, r_max=16, n_r=64, n_angles=5, num_turns=300, nemitt_x=2.5e-6,
                     nemitt_y=2.5e-6)
    assert len(res['da']) == 5
    for (th, d), v in res['da'].items():
        ana = 0.02 / np.sqrt(geps * np.max(np.cos(th) ** 2 * bx + np.sin(th) ** 2 * by))
        assert ana - step < v <= ana + 1e-9, (th, v, ana)
    # below the aperture every particle survives; the survival turn of the
    # lost ones is the turn count at loss
    surv = res['survival']
    assert surv.max() == 300 and surv.min() < 300
    assert np.all(surv[res['r'] <= 9.0] == 300)


def test_normalized_particles_follow_the_normal_form():
    """The inverse normal form returns the normalised input, each plane
    scaled by its own emittance, at a momentum deviation other than 0."""
    from xtrack_amd.twiss import (find_closed_orbit, compute_one_turn_matrix_finite_differences,
                                  get_linear_normal_form)
    line = lattices.thin_fodo_ring()
    line.build_tracker(backend=OracleBackend())
    ref = line.particle_ref
    rng = np.random.default_rng(3)
    xn, pxn, yn, pyn = rng.normal(0, 2, (4, 20))
    p = da.build_normalized_particles(line, xn, pxn, yn, pyn, 1e-3, 2e-6, 3e-6)
    geps = np.array([2e-6, 3e-6]) / float(ref._np('beta0')[0] * ref._np('gamma0')[0])
    co = find_closed_orbit(line, ref, method='4d', delta0=1e-3)
    R, _ = compute_one_turn_matrix_finite_differences(line, ref, co, kill_cavity=True)
    W, invW, _, _ = get_linear_normal_form(R, only_4d_block=True)
    h = p.to_numpy()
    z = np.stack([h['x'], h['px'], h['y'], h['py']]) - co[:4, None]
    back = invW[:4, :4] @ z
    assert np.allclose(back[0] / np.sqrt(geps[0]), xn, rtol=1e-9, atol=1e-9)
    assert np.allclose(back[1] / np.sqrt(geps[0]), pxn, rtol=1e-9, atol=1e-9)
    assert np.allclose(back[2] / np.sqrt(geps[1]), yn, rtol=1e-9, atol=1e-9)
    assert np.allclose(back[3] / np.sqrt(geps[1]), pyn, rtol=1e-9, atol=1e-9)
    assert np.all(h['delta'] == 1e-3)


def test_polar_grid_and_aperture_extraction():
    """Grid layout, and the aperture as the last amplitude below the first
    loss at each angle (r_max when nothing is lost, 0 when the first point
    is lost), on synthetic survival turns."""
    r, th, d = da.polar_grid(10.0, 10, 4, deltas=(0.0, 1e-3))
    assert len(r) == 10 * 4 * 2 and np.isclose(r.min(), 1.0) and np.isclose(r.max(), 10.0)
    assert np.allclose(np.unique(th), (np.arange(4) + 0.5) * np.pi / 8)
    T = 100
    surv = np.full(len(r), T)
    a0, a1 = np.unique(th)[:2]
    surv[(th == a0) & (d == 0.0) & (r >= 6.0)] = 10                   # lost from 6 sigma
    surv[(th == a1) & (d == 0.0) & (np.isclose(r, 3.0))] = 50         # isolated loss at 3
    surv[(th == a0) & (d == 1e-3)] = 1                                 # all lost
    ap = da.dynamic_aperture(r, th, d, surv, T)
    assert np.isclose(ap[(float(a0), 0.0)], 5.0)
    assert np.isclose(ap[(float(a1), 0.0)], 2.0)
    assert ap[(float(a0), 1e-3)] == 0.0
    assert np.isclose(ap[(float(np.unique(th)[3]), 0.0)], 10.0)
