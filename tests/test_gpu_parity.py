"""GPU (libxtrack_amd.so) against the CPU oracle built with the same
elementary-function library: every coordinate and every loss record must be
identical (IEEE equality)."""
import numpy as np
import pytest

import xtrack_amd as xt
from helpers import (mixed_line, segment_line, random_particles, compare_exact,
                     guard_edge_particles)
from oracle_backend import OracleBackend

pytestmark = pytest.mark.gpu

_MODE = {'compile': False}


@pytest.fixture(params=[False, True], ids=['interpreter', 'compiled'], autouse=True)
def _track_mode(request):
    """Every test runs on the element-program interpreter and on the
    line's compiled code."""
    _MODE['compile'] = request.param
    yield


def _pair(line_fn, p):
    line_g = line_fn()
    line_g.build_tracker(device='cuda:0', compile=_MODE['compile'])
    assert line_g.tracker.backend.is_compiled(line_g.tracker._handle) == _MODE['compile']
    line_c = line_fn()
    line_c.build_tracker(backend=OracleBackend())
    pg = p.copy(_context='cuda:0')
    pc = p.copy()
    return line_g, line_c, pg, pc


@pytest.mark.parametrize('n_turns', [1, 20])
def test_mixed_ring_bit_exact(n_turns):
    print(f'[mixed ring] n_turns={n_turns}', flush=True)
    p = random_particles(2000, seed=n_turns, scale=4e-3)
    lg, lc, pg, pc = _pair(lambda: mixed_line(n_cells=3), p)
    lg.track(pg, num_turns=n_turns)
    lc.track(pc, num_turns=n_turns)
    bad = compare_exact(pg, pc)
    n_lost = int((pc.sorted_by_id()['state'] <= 0).sum())
    print(f'lost={n_lost}', flush=True)
    assert not bad, bad


def test_reciprocal_division_range_edges():
    """Lanes on both sides of the reciprocal-division ranges, mixed within
    waves with ordinary lanes."""
    print('[division range edges]', flush=True)
    for seed, kw in ((4, dict(num_turns=2)), (5, dict(num_elements=7))):
        p = guard_edge_particles(3000, seed=seed)
        lg, lc, pg, pc = _pair(lambda: mixed_line(n_cells=2, aperture=0.95), p)
        lg.track(pg, **kw)
        lc.track(pc, **kw)
        bad = compare_exact(pg, pc)
        assert not bad, bad


def test_losses_on_every_element_kind():
    print('[losses]', flush=True)
    p = random_particles(4000, seed=7, scale=8e-3, spread=1.5)
    lg, lc, pg, pc = _pair(lambda: mixed_line(n_cells=2, aperture=0.012), p)
    lg.track(pg, num_turns=30)
    lc.track(pc, num_turns=30)
    bad = compare_exact(pg, pc)
    sc = pc.sorted_by_id()
    assert (sc['state'] <= 0).sum() > 100
    assert len(np.unique(sc['at_element'][sc['state'] <= 0])) > 3
    assert not bad, bad


def test_ions_chi_not_one_literal_path():
    """chi != 1 takes the literal kick on every lane."""
    print('[ions]', flush=True)
    rng = np.random.default_rng(11)
    n = 1500
    p = xt.Particles(p0c=7e9, x=rng.normal(0, 2e-3, n), px=rng.normal(0, 2e-4, n),
                     y=rng.normal(0, 2e-3, n), py=rng.normal(0, 2e-4, n),
                     zeta=rng.normal(0, 0.05, n), delta=rng.normal(0, 1e-3, n),
                     charge_ratio=rng.choice([0.5, 1.0, 2.0], n),
                     mass_ratio=rng.choice([1.0, 2.0, 4.0], n))
    lg, lc, pg, pc = _pair(lambda: mixed_line(n_cells=2), p)
    lg.track(pg, num_turns=10)
    lc.track(pc, num_turns=10)
    assert not compare_exact(pg, pc)


def test_nonfinite_and_extreme_particles():
    """NaN / inf / huge coordinates: identical results and loss records."""
    print('[extreme]', flush=True)
    rng = np.random.default_rng(12)
    n = 600
    x = rng.normal(0, 2e-3, n)
    px = rng.normal(0, 2e-4, n)
    x[:50] = np.nan
    px[50:100] = np.inf
    px[100:150] = rng.uniform(0.5, 3.0, 50)
    x[150:200] = rng.uniform(-0.99, 0.99, 50)
    p = xt.Particles(p0c=7e9, x=x, px=px, y=rng.normal(0, 1e-3, n), delta=rng.normal(0, 1e-3, n))
    lg, lc, pg, pc = _pair(lambda: mixed_line(n_cells=2, aperture=0.5), p)
    lg.track(pg, num_turns=3)
    lc.track(pc, num_turns=3)
    assert not compare_exact(pg, pc)
    s = pc.sorted_by_id()
    assert set(np.unique(s['state'])) >= {-1, 0}


def test_global_aperture_after_static_thick_elements():
    """xtrack tests/test_apertures.py:16-43 on the GPU."""
    def mk():
        return xt.Line(elements=[xt.Drift(length=2), xt.Quadrupole(length=2, k1=0),
                                 xt.Multipole(length=2, isthick=True)])
    line = mk()
    line.build_tracker(device='cuda:0', compile=_MODE['compile'])
    for start, expected in ((1, -1), (2, 1), (0, -1)):
        p = xt.Particles(p0c=7e12, px=0.6)
        p.move(_context='cuda:0')
        line.track(p, ele_start=start, num_elements=1)
        assert int(p.state.cpu()[0]) == expected


def test_partial_tracking_matrix():
    """tests/test_tracker.py:152-340: every (start, stop, turns) split."""
    n_elem = 9
    line = xt.Line(elements=[xt.Drift(length=1.) for _ in range(n_elem)])
    line.build_tracker(device='cuda:0', compile=_MODE['compile'])
    p0 = xt.Particles(x=[1e-3, -2e-3, 5e-3], y=[2e-3, -4e-3, 3e-3], zeta=1e-2, p0c=7e12)

    def run(expected_turn, expected_el, **kw):
        p = p0.copy(_context='cuda:0')
        if 'ele_start' in kw:
            p.at_element = kw['ele_start']
            p.s = float(kw['ele_start'])
        line.track(p, turn_by_turn_monitor=True, **kw)
        h = p.to_numpy()
        assert np.all(h['at_turn'] == expected_turn), kw
        assert np.all(h['at_element'] == expected_el), kw
        assert np.all(h['s'] == expected_el), kw
        nmon = expected_turn if expected_el == 0 else expected_turn + 1
        assert line.record_last_track.x.shape == (3, nmon), kw

    for turns in (1, 2, 10):
        run(turns, 0, num_turns=turns)
        for start in range(n_elem):
            run(turns, 0, num_turns=turns, ele_start=start)
        for stop in range(1, n_elem):
            run(turns - 1, stop, num_turns=turns, ele_stop=stop)
        for start in range(n_elem):
            for stop in range(start + 1, n_elem):
                run(turns - 1, stop, num_turns=turns, ele_start=start, ele_stop=stop)
            for stop in range(start + 1):
                run(turns, stop, num_turns=turns, ele_start=start, ele_stop=stop)
    for start in range(n_elem):
        for shift in range(1, 3 * n_elem + 1):
            t = (start + shift) // n_elem
            run(t, start + shift - n_elem * t, ele_start=start, num_elements=shift)


@pytest.mark.parametrize('ele_start,ele_stop,expected_x', [
    (None, None, [0, 0.005, 0.010, 0.015, 0.020, 0.025]),
    (None, 3, [0, 0.005, 0.010, 0.015, 0.020, 0.023]),
    (2, None, [0, 0.003, 0.008, 0.013, 0.018, 0.023]),
    (2, 3, [0, 0.003, 0.008, 0.013, 0.018, 0.021]),
    (3, 2, [0, 0.002, 0.007, 0.012, 0.017, 0.022, 0.024]),
])
def test_tbt_monitor_series(ele_start, ele_stop, expected_x):
    """Expected series of xtrack tests/test_tracker.py:577-602."""
    line = xt.Line(elements=[xt.Drift(length=1)] * 5)
    line.build_tracker(device='cuda:0', compile=_MODE['compile'])
    p = xt.Particles(px=0.001)
    p.move(_context='cuda:0')
    line.track(p, num_turns=5, turn_by_turn_monitor=True, ele_start=ele_start,
               ele_stop=ele_stop)
    rec = line.record_last_track.x
    assert rec.shape == (1, len(expected_x) - 1)
    got = np.concatenate([rec[0], p.x.cpu().numpy()])
    assert np.allclose(got, expected_x, atol=1e-16, rtol=0)


def test_monitors_bit_exact():
    """Turn-by-turn and element-by-element records, GPU vs oracle."""
    p = random_particles(300, seed=5, scale=3e-3)
    for mode in (True, 'ONE_TURN_EBE'):
        lg, lc, pg, pc = _pair(lambda: mixed_line(n_cells=2), p)
        nt = 6 if mode is True else 1
        lg.track(pg, num_turns=nt, turn_by_turn_monitor=mode)
        lc.track(pc, num_turns=nt, turn_by_turn_monitor=mode)
        mg, mc = lg.record_last_track, lc.record_last_track
        for ff in ('x', 'px', 'y', 'py', 'zeta', 'delta', 'at_turn', 'at_element', 'state',
                   'particle_id'):
            a, b = getattr(mg, ff), getattr(mc, ff)
            assert a.shape == b.shape
            assert np.array_equal(a, b, equal_nan=True), (mode, ff)


@pytest.mark.parametrize('kw', [dict(num_turns=2), dict(ele_start=3, ele_stop=11),
                                dict(num_turns=1, turn_by_turn_monitor='ONE_TURN_EBE')])
def test_backtracking_bit_exact(kw):
    """Forward two turns, then backtrack (the interpreter runs the
    backtracking programs whichever kernel tracks forward)."""
    print(f'[backtrack] {kw}', flush=True)
    p = random_particles(1500, seed=8, scale=5e-3)
    lg, lc, pg, pc = _pair(lambda: mixed_line(n_cells=2), p)
    lg.track(pg, num_turns=2)
    lc.track(pc, num_turns=2)
    lg.track(pg, backtrack=True, **kw)
    lc.track(pc, backtrack=True, **kw)
    bad = compare_exact(pg, pc)
    assert not bad, bad
    if 'turn_by_turn_monitor' in kw:
        for ff in ('x', 'px', 'zeta', 'at_element', 'at_turn', 'state'):
            assert np.array_equal(np.asarray(getattr(lg.record_last_track, ff)),
                                  getattr(lc.record_last_track, ff), equal_nan=True), ff


def test_backtracking_exact_bends_bit_exact():
    """Backtracking through the exact bend-kick-bend maps (curved and
    straight frames, with gradient and multipole kicks)."""
    from test_exact_bend import _bkb_ring
    print('[backtrack exact bends]', flush=True)
    p = random_particles(1500, seed=18, scale=4e-3)
    lg, lc, pg, pc = _pair(_bkb_ring, p)
    for ll, pp in ((lg, pg), (lc, pc)):
        ll.track(pp, num_turns=3)
        ll.track(pp, num_turns=2, backtrack=True)
    bad = compare_exact(pg, pc)
    assert not bad, bad


def test_sps_lattice_bit_exact():
    """The SPS line of the reference's test data (JSON: Drift, DipoleEdge,
    Multipole, apertures, one shifted LimitRect, cavities) with losses."""
    from test_sps_lattice import sps_line, sps_particles
    print('[sps lattice]', flush=True)
    p = sps_particles(3000, seed=1)
    lg, lc, pg, pc = _pair(sps_line, p)
    lg.track(pg, num_turns=3)
    lc.track(pc, num_turns=3)
    bad = compare_exact(pg, pc)
    sc = pc.sorted_by_id()
    assert (sc['state'] <= 0).sum() > 200
    assert not bad, bad
    lg.track(pg, num_turns=1, backtrack=True)
    lc.track(pc, num_turns=1, backtrack=True)
    bad = compare_exact(pg, pc)
    assert not bad, bad


def test_misaligned_elements_bit_exact():
    """Shifts, rolls, pitch and yaw with anchors on every class that allows
    them, straight and curved frames, forward and backtracking."""
    from test_misalignments import _misaligned_line
    print('[misaligned elements]', flush=True)
    p = random_particles(3000, seed=12, scale=5e-3)
    for bt in (False, True):
        lg, lc, pg, pc = _pair(lambda: xt.Line(elements=_misaligned_line()), p)
        lg.track(pg, num_turns=4, backtrack=bt)
        lc.track(pc, num_turns=4, backtrack=bt)
        bad = compare_exact(pg, pc)
        assert (pc.state <= 0).sum() > 100
        assert not bad, (bt, bad)


def test_rf_multipoles_bit_exact():
    from test_rf_multipole import rf_line
    print('[rf multipoles]', flush=True)
    rng = np.random.default_rng(21)
    n = 2000
    p = xt.Particles(p0c=7e9, x=rng.normal(0, 4e-3, n), px=rng.normal(0, 1e-4, n),
                     y=rng.normal(0, 4e-3, n), py=rng.normal(0, 1e-4, n),
                     zeta=rng.normal(0, 0.2, n), delta=rng.normal(0, 1e-3, n),
                     charge_ratio=rng.choice([1.0, 0.5], n), mass_ratio=rng.choice([1.0, 2.0], n))
    for bt in (False, True):
        lg, lc, pg, pc = _pair(lambda: xt.Line(elements=rf_line()), p)
        lg.track(pg, num_turns=3, backtrack=bt)
        lc.track(pc, num_turns=3, backtrack=bt)
        bad = compare_exact(pg, pc)
        assert not bad, (bt, bad)


def test_hllhc_lattice_bit_exact():
    """HL-LHC v1.4 of the reference's test data (beam-beam elements
    replaced by markers): thin multipoles, crab cavities, cavities."""
    from test_hllhc_lattice import hllhc_line, hllhc_particles
    print('[hllhc lattice]', flush=True)
    p = hllhc_particles(3000, seed=3)
    lg, lc, pg, pc = _pair(hllhc_line, p)
    lg.track(pg, num_turns=3)
    lc.track(pc, num_turns=3)
    bad = compare_exact(pg, pc)
    assert (pc.sorted_by_id()['state'] <= 0).sum() > 100
    assert not bad, bad


def test_twiss_through_gpu_tracker():
    """Twiss (closed orbit, one-turn matrix, normal form, element-by-element
    propagation) with the GPU tracker equals the one with the oracle: the
    probes are tracked bit for bit."""
    from xtrack_amd import lattices
    print('[twiss]', flush=True)

    def mk():
        return lattices.collider_ring(n_arcs=1, n_cells=4, k2f=0.03, k2d=-0.06)

    lg, lc = mk(), mk()
    lg.build_tracker(device='cuda:0', compile=_MODE['compile'])
    lc.build_tracker(backend=OracleBackend())
    for method in ('4d', '6d'):
        tg, tc = lg.twiss(method=method), lc.twiss(method=method)
        for nn in ('betx', 'bety', 'alfx', 'alfy', 'mux', 'muy', 'dx', 'dpx', 'x', 'px',
                   'zeta', 'pzeta'):
            assert np.array_equal(tg[nn], tc[nn]), (method, nn)
        assert tg.qx == tc.qx and tg.qy == tc.qy and tg.dqx == tc.dqx
        assert np.array_equal(tg.R_matrix, tc.R_matrix)


def test_sliced_line_bit_exact():
    """Thin, thick, drift and edge slices (misaligned parents included),
    forward and backtracking."""
    from test_slicing import sliced_line
    print('[sliced line]', flush=True)
    p = random_particles(3000, seed=14, scale=4e-3)
    for bt in (False, True):
        lg, lc, pg, pc = _pair(sliced_line, p)
        lg.track(pg, num_turns=3, backtrack=bt)
        lc.track(pc, num_turns=3, backtrack=bt)
        bad = compare_exact(pg, pc)
        assert not bad, (bt, bad)


def test_segment_runs_bit_exact():
    """Runs of cheap elements that the compiled line code tracks as
    segments: losses at every aperture inside the runs, tracking ranges
    that start or stop inside a segment, and element-by-element records."""
    print('[segment runs]', flush=True)
    p = random_particles(3000, seed=1, scale=2e-3)
    for kw in (dict(num_turns=3), dict(num_turns=2, ele_start=5, ele_stop=50),
               dict(ele_start=20, num_elements=70), dict(ele_start=16, ele_stop=27),
               dict(num_turns=1, turn_by_turn_monitor='ONE_TURN_EBE')):
        print(kw, flush=True)
        lg, lc, pg, pc = _pair(segment_line, p)
        lg.track(pg, **kw)
        lc.track(pc, **kw)
        bad = compare_exact(pg, pc)
        assert not bad, (kw, bad)
        if 'turn_by_turn_monitor' in kw:
            for ff in ('x', 'px', 'y', 'py', 'zeta', 'at_element', 'state'):
                assert np.array_equal(getattr(lg.record_last_track, ff),
                                      getattr(lc.record_last_track, ff), equal_nan=True), ff


def test_da_scan_survival_turns():
    """A dynamic-aperture scan on a sextupole-perturbed ring: the GPU
    tracker gives the oracle's survival turn for every grid particle."""
    from xtrack_amd import lattices, da
    print('[da scan]', flush=True)

    ring = lattices.sextupole_fodo_ring

    lg, lc = ring(), ring()
    lg.build_tracker(device='cuda:0', compile=_MODE['compile'])
    lc.build_tracker(backend=OracleBackend())
    kw = dict(r_max=12, n_r=24, n_angles=5, num_turns=200, nemitt_x=2.5e-6, nemitt_y=2.5e-6,
              deltas=(0.0, 2e-3))
    rg, rc = da.da_scan(lg, **kw), da.da_scan(lc, **kw)
    assert np.array_equal(rg['survival'], rc['survival'])
    assert rg['da'] == rc['da']
    surv = rc['survival']
    assert 0 < (surv < 200).sum() < len(surv)
    assert len(np.unique(surv)) > 5


def test_freeze_longitudinal_bit_exact():
    """Line.track(freeze_longitudinal=True): the kernel (energy kicks
    skipped, zeta restored after every element) against the oracle (the
    reference's frozen setters)."""
    print('[freeze longitudinal]', flush=True)
    p = random_particles(2000, seed=21, scale=4e-3)
    lg, lc, pg, pc = _pair(lambda: mixed_line(n_cells=3), p)
    lg.track(pg, num_turns=5, freeze_longitudinal=True)
    lc.track(pc, num_turns=5, freeze_longitudinal=True)
    bad = compare_exact(pg, pc)
    assert not bad, bad
    assert np.array_equal(pg.sorted_by_id()['zeta'], p.sorted_by_id()['zeta'])


def test_exact_bend_bit_exact():
    """'bend-kick-bend' bends (exact curved and straight dipole maps, the
    small-field fallbacks, with kicks of k1 and multipoles), protons and
    ions with chi != 1."""
    from test_exact_bend import _bkb_ring, _ions
    print('[exact bend]', flush=True)
    for p in (random_particles(3000, seed=22, scale=4e-3), _ions(2000, seed=23)):
        lg, lc, pg, pc = _pair(_bkb_ring, p)
        lg.track(pg, num_turns=6)
        lc.track(pc, num_turns=6)
        bad = compare_exact(pg, pc)
        assert not bad, bad
