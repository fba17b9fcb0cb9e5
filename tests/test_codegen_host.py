"""Generated line code (xtm_line_specialize's source) compiled for the host
against the CPU oracle: identical bits for every coordinate and loss
record, on the same cases as the interpreter's host tests."""
import numpy as np
import pytest

import xtrack_amd as xt
from helpers import mixed_line, random_particles, compare_exact, guard_edge_particles
from oracle_backend import OracleBackend, GeneratedHostBackend, generated_host_line
from xtrack_amd.elements import build_raw_table


def _run(line_fn, p, **kw):
    lg = line_fn()
    lg.build_tracker(backend=GeneratedHostBackend())
    lc = line_fn()
    lc.build_tracker(backend=OracleBackend())
    pg, pc = p.copy(), p.copy()
    lg.track(pg, **kw)
    lc.track(pc, **kw)
    return lg, lc, pg, pc


@pytest.mark.parametrize('n_turns', [1, 12])
def test_mixed_ring(n_turns):
    p = random_particles(600, seed=n_turns, scale=4e-3)
    _, _, pg, pc = _run(lambda: mixed_line(n_cells=3), p, num_turns=n_turns)
    assert not compare_exact(pg, pc)


def test_ions_and_range_edges():
    rng = np.random.default_rng(11)
    n = 300
    p = xt.Particles(p0c=7e9, x=rng.normal(0, 2e-3, n), px=rng.normal(0, 2e-4, n),
                     y=rng.normal(0, 2e-3, n), py=rng.normal(0, 2e-4, n),
                     zeta=rng.normal(0, 0.05, n), delta=rng.normal(0, 1e-3, n),
                     charge_ratio=rng.choice([0.5, 1.0, 2.0], n),
                     mass_ratio=rng.choice([1.0, 2.0, 4.0], n))
    _, _, pg, pc = _run(lambda: mixed_line(n_cells=2), p, num_turns=5)
    assert not compare_exact(pg, pc)
    p = guard_edge_particles(1200, seed=4)
    _, _, pg, pc = _run(lambda: mixed_line(n_cells=2, aperture=0.95), p, num_turns=2)
    assert not compare_exact(pg, pc)


def test_partial_tracking_and_monitors():
    p = random_particles(200, seed=5, scale=3e-3)
    _, _, pg, pc = _run(lambda: mixed_line(n_cells=2), p, num_turns=3, ele_start=5,
                        ele_stop=17)
    assert not compare_exact(pg, pc)
    for mode, nt in ((True, 4), ('ONE_TURN_EBE', 1)):
        lg, lc, pg, pc = _run(lambda: mixed_line(n_cells=2), p, num_turns=nt,
                              turn_by_turn_monitor=mode)
        for ff in ('x', 'px', 'zeta', 'delta', 'at_turn', 'at_element', 'state'):
            assert np.array_equal(getattr(lg.record_last_track, ff),
                                  getattr(lc.record_last_track, ff), equal_nan=True), ff


def test_collider_cells():
    from xtrack_amd import lattices
    p = random_particles(64, seed=3, scale=5e-4, p0c=7e12, spread=0.3)
    _, _, pg, pc = _run(lambda: lattices.collider_ring(n_arcs=1, n_cells=2, k2f=0.03,
                                                       k2d=-0.06), p, num_turns=3)
    assert not compare_exact(pg, pc)


def test_source_shape():
    """One function per distinct record; elements map onto them."""
    from xtrack_amd import lattices
    line = lattices.collider_ring(k2f=0.03, k2d=-0.06)
    t, o, pp = build_raw_table(line.elements)
    _, rec, prog, src, _ = generated_host_line(t, o, pp, line.get_length())
    assert len(np.unique(rec)) == 10
    assert src.count('XT_GEN_FN void rec_') == 10
    assert 'xt::op_yoshida4(p' in src and 'xt::op_limit_rect(p' in src
