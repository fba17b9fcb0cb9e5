"""Host logic of the screened dynamic-aperture scan: subset layouts, the
per-angle interpolation used as the no-network predictor, and the boundary
band, on a synthetic grid whose survival turns are known in closed form."""
import types

import numpy as np

from xtrack_amd import da, surrogate as S


def _grid(n_lines=2, r_max=10.0, n_r=40, n_angles=3, deltas=(0.0, 1e-3)):
    """GridSet's host attributes without particles; n_r per lattice."""
    gs = types.SimpleNamespace()
    n_r = [n_r] * n_lines if np.ndim(n_r) == 0 else list(n_r)
    grids = [da.polar_grid(r_max, nr, n_angles, deltas) for nr in n_r]
    gs.r_slot = np.concatenate([g[0] for g in grids])
    gs.theta_slot = np.concatenate([g[1] for g in grids])
    gs.delta_slot = np.concatenate([g[2] for g in grids])
    gs.r_max, gs.n_lines = r_max, n_lines
    gs.starts = np.concatenate([[0], np.cumsum([len(g[0]) for g in grids])])
    gs.n_slots = int(gs.starts[-1])
    gs.line_of = lambda slots: np.searchsorted(gs.starts, slots, side='right') - 1
    return gs


def _truth(gs, T, da_of):
    """Survival: T below the aperture da_of(line, theta), falling as
    T * exp(-(r - da)) above it."""
    slots = np.arange(gs.n_slots)
    ap = np.array([da_of(l, th) for l, th in zip(gs.line_of(slots), gs.theta_slot)])
    r = gs.r_slot
    return np.where(r <= ap, T, np.floor(T * np.exp(-(r - ap)))).astype(np.int64), ap


def test_subset_layout():
    starts = np.array([0, 100, 250, 300])
    slots = np.array([3, 50, 99, 100, 180, 260, 299])
    assert list(S._starts_of(slots, starts)) == [0, 3, 5, 7]
    assert list(S._starts_of(np.array([], dtype=np.int64), starts)) == [0, 0, 0, 0]


def test_interp_predictor_and_band_find_the_aperture():
    T = 1000
    gs = _grid(n_r=(40, 55))
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
    step = gs.r_slot[1] - gs.r_slot[0]
    first_loss = np.abs(gs.r_slot[rest] - ap[rest] - step) < 0.75 * step
    assert np.all(np.isin(rest[first_loss & (gs.r_slot[rest] < gs.r_max)], band))
    assert len(band) < 0.5 * len(rest)


def test_band_factor_selects_predicted_survival_near_target():
    T = 1000
    gs = _grid(n_lines=1)
    guess = np.full(gs.n_slots, float(T))
    rest = np.arange(gs.n_slots)
    guess[:10] = [1, 10, 200, 333, 334, 1000, 2999, 3001, 5, 500]
    band = S.boundary_band(gs, guess, rest, T, band_factor=3.0, band_sigma=-1.0)
    assert set(band[band < 10].tolist()) == {4, 5, 6, 9}


def test_predicted_aperture_uses_half_the_turns():
    """A group whose predictions sit just below num_turns (survivors as the
    network sees them) has no predicted loss: the band lies at the top of
    the grid, not at its first particle."""
    T = 1000
    gs = _grid(n_lines=1, deltas=(0.0,), n_angles=1)
    guess = np.full(gs.n_slots, 0.9 * T)
    rest = np.arange(gs.n_slots)
    band = S.boundary_band(gs, guess, rest, T, band_factor=1.0, band_sigma=0.5)
    assert len(band) and gs.r_slot[band].min() >= gs.r_max - 0.5 - 1e-9
    guess[30:] = 0.2 * T
    band = S.boundary_band(gs, guess, rest, T, band_factor=1.0, band_sigma=0.5)
    r = gs.r_slot
    assert np.allclose(sorted(r[band]), sorted(r[np.abs(r - r[30]) <= 0.5]))


def test_below_first_loss_marks_the_slots_that_can_change_the_aperture():
    """Per (lattice, angle, delta) group: the slots below the smallest lost
    amplitude; a group without losses keeps every slot; the aperture of the
    survival turns is unchanged when only those slots are known."""
    T = 1000
    gs = _grid(n_r=(40, 55))
    da_of = lambda l, th: 4.0 + l + 2 * th
    surv, ap = _truth(gs, T, da_of)
    lost = surv < T
    below = S.below_first_loss(gs, lost)
    assert not np.any(below & lost)
    assert np.all(gs.r_slot[below] <= ap[below] + 1e-12)
    # slots above the first loss carry no information on the aperture
    masked = surv.copy()
    masked[~below & ~lost] = 1
    for l in range(gs.n_lines):
        sl = slice(gs.starts[l], gs.starts[l + 1])
        assert da.dynamic_aperture(gs.r_slot[sl], gs.theta_slot[sl], gs.delta_slot[sl],
                                   masked[sl], T) == \
            da.dynamic_aperture(gs.r_slot[sl], gs.theta_slot[sl], gs.delta_slot[sl],
                                surv[sl], T)
    # no loss at all: every slot counts
    assert np.all(S.below_first_loss(gs, np.zeros(gs.n_slots, bool)))


def test_band_uses_the_ensemble_range():
    """A slot is in the factor band when any ensemble member's prediction
    falls near the target; the predicted aperture uses the smallest
    prediction."""
    T = 1000
    gs = _grid(n_lines=1, deltas=(0.0,), n_angles=1)
    rest = np.arange(gs.n_slots)
    guess = np.full(gs.n_slots, 10.0 * T)
    lo, hi = guess.copy(), guess.copy()
    lo[5], hi[5] = 2000.0, 8000.0        # one member within a factor 3
    lo[7], hi[7] = 5000.0, 9000.0        # none
    band = S.boundary_band(gs, guess, rest, T, 3.0, -1.0, guess_lo=lo, guess_hi=hi)
    assert 5 in band and 7 not in band
    lo[20] = 100.0                       # one member predicts a loss at slot 20
    band = S.boundary_band(gs, guess, rest, T, 1.0, 0.26, guess_lo=lo, guess_hi=hi)
    assert set(band.tolist()) == {k for k in rest if abs(gs.r_slot[k] - gs.r_slot[20]) <= 0.26}


def test_small_sum_equals_the_plain_sum():
    import torch
    g = torch.Generator().manual_seed(0)
    for shape in ((1024, 256), (1000, 7), (5,), (1, 3), (130,), (20000, 2), (128,)):
        a = torch.randn(shape, generator=g, dtype=torch.float64)
        assert torch.allclose(S._small_sum(a), a.sum(dim=0), rtol=0, atol=1e-12)


def test_features_tune_diffusion_from_accumulators():
    """Synthetic accumulators of three particles over a 10-turn window:
    constant phase advance 0.31 turn (no diffusion), advance 0.30 in the
    first half and 0.32 in the second (diffusion 0.02), and a particle lost
    after 4 turns (no second half: diffusion input 0)."""
    import torch
    window = 10
    acc = torch.zeros((17, 3), dtype=torch.float64)
    two_pi = 2 * np.pi
    counts = [10, 10, 4]
    adv1 = [0.31, 0.30, 0.25]
    adv2 = [0.31, 0.32, 0.25]
    for i in range(3):
        n = counts[i]
        acc[0, i] = n
        # advances at turns 1..n-1; second half: turns >= 5
        turns = np.arange(1, n)
        ph = np.where(turns >= 5, adv2[i], adv1[i]) * two_pi
        acc[7, i] = acc[8, i] = ph.sum()
        acc[15, i] = acc[16, i] = ph[turns >= 5].sum()
        acc[1, i] = acc[4, i] = n * 2.0          # mean action 2
        acc[2, i] = acc[5, i] = n * 4.0
        acc[3, i] = acc[6, i] = 2.0
        acc[13, i] = acc[14, i] = 2.0
    f = S.features_from_acc(acc, np.array([1.0, 1.0, 1.0]), np.zeros(3), np.zeros(3), window)
    f = f.numpy().astype(np.float64)
    # tunes: mean advance per turn over the window
    assert np.allclose(f[:, 11], [0.31, (4 * 0.30 + 5 * 0.32) / 9, 0.25], atol=1e-6)
    # log10 tune diffusion: 0 -> -12 (floor), 0.02, and 0 for the lost one
    assert f[0, 14] < -6 and abs(f[1, 14] - np.log10(0.02)) < 1e-4 and f[2, 14] == 0.0
    assert np.array_equal(f[:, 14], f[:, 15])
    # alive fraction, action features
    assert np.allclose(f[:, 13], [1.0, 1.0, 0.4])
    assert np.allclose(f[:, 5], 0.0) and np.allclose(f[:, 7], 0.0)
