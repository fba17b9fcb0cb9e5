"""Host logic of the screened dynamic-aperture scan: subset layouts, the
per-angle interpolation used as the no-network predictor, and the boundary
band, on a synthetic grid whose survival turns are known in closed form."""
import types

import numpy as np

from xtrack_amd import da, surrogate as S


def _grid(n_lines=2, r_max=10.0, n_r=40, n_angles=3, deltas=(0.0, 1e-3)):
    gs = types.SimpleNamespace()
    gs.r, gs.theta, gs.delta = da.polar_grid(r_max, n_r, n_angles, deltas)
    gs.n_grid, gs.n_lines = len(gs.r), n_lines
    gs.starts = np.arange(n_lines + 1) * gs.n_grid
    gs.line_of = lambda slots: np.searchsorted(gs.starts, slots, side='right') - 1
    return gs


def _truth(gs, T, da_of):
    """Survival: T below the aperture da_of(line, theta), falling as
    T * exp(-(r - da)) above it."""
    slots = np.arange(gs.n_grid * gs.n_lines)
    g = slots % gs.n_grid
    ap = np.array([da_of(l, th) for l, th in zip(gs.line_of(slots), gs.theta[g])])
    r = gs.r[g]
    return np.where(r <= ap, T, np.floor(T * np.exp(-(r - ap)))).astype(np.int64), ap


def test_subset_layout():
    starts = np.array([0, 100, 250, 300])
    slots = np.array([3, 50, 99, 100, 180, 260, 299])
    assert list(S._starts_of(slots, starts)) == [0, 3, 5, 7]
    assert list(S._starts_of(np.array([], dtype=np.int64), starts)) == [0, 0, 0, 0]


def test_interp_predictor_and_band_find_the_aperture():
    T = 1000
    gs = _grid()
    da_of = lambda l, th: 4.0 + l + 2 * th
    surv, ap = _truth(gs, T, da_of)
    slots = np.arange(len(surv))
    rng = np.random.default_rng(0)
    lab = np.sort(rng.choice(slots, size=len(slots) // 4, replace=False))
    rest = np.setdiff1d(slots, lab)
    pred = S.interp_predictor(gs, lab, surv, rest)
    # exact on the labels' own amplitudes, between labels bounded by them
    assert np.all(pred > 0)
    guess = surv.astype(float)
    guess[rest] = pred
    band = S.boundary_band(gs, guess, rest, T, band_factor=1.0, band_sigma=1.0)
    # every particle of `rest` at the true aperture's first loss is in the band
    g = rest % gs.n_grid
    first_loss = np.abs(gs.r[g] - ap[rest] - (gs.r[1] - gs.r[0])) < 0.75 * (gs.r[1] - gs.r[0])
    assert np.all(np.isin(rest[first_loss & (gs.r[g] < gs.r.max())], band))
    assert len(band) < 0.5 * len(rest)


def test_band_factor_selects_predicted_survival_near_target():
    T = 1000
    gs = _grid(n_lines=1)
    guess = np.full(gs.n_grid, float(T))
    rest = np.arange(gs.n_grid)
    guess[:10] = [1, 10, 200, 333, 334, 1000, 2999, 3001, 5, 500]
    band = S.boundary_band(gs, guess, rest, T, band_factor=3.0, band_sigma=-1.0)
    assert set(band[band < 10].tolist()) == {4, 5, 6, 9}


def test_predicted_aperture_uses_half_the_turns():
    """A group whose predictions sit just below num_turns (survivors as the
    network sees them) has no predicted loss: the band lies at the top of
    the grid, not at its first particle."""
    T = 1000
    gs = _grid(n_lines=1, deltas=(0.0,), n_angles=1)
    guess = np.full(gs.n_grid, 0.9 * T)
    rest = np.arange(gs.n_grid)
    band = S.boundary_band(gs, guess, rest, T, band_factor=1.0, band_sigma=0.5)
    assert len(band) and gs.r[band].min() >= gs.r.max() - 0.5 - 1e-9
    guess[30:] = 0.2 * T
    band = S.boundary_band(gs, guess, rest, T, band_factor=1.0, band_sigma=0.5)
    assert np.allclose(sorted(gs.r[band]), sorted(gs.r[np.abs(gs.r - gs.r[30]) <= 0.5]))
