"""Generated line code (xtm_line_specialize's source) compiled for the host
against the CPU oracle: identical bits for every coordinate and loss
record, on the same cases as the interpreter's host tests."""
import numpy as np
import pytest

import xtrack_amd as xt
from helpers import (mixed_line, segment_line, random_particles, compare_exact,
                     guard_edge_particles)
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
    """One function per distinct record; elements map onto them.  Runs of
    cheap records (drifts, apertures, thin kicks) start with a key that
    names a segment function tracking the whole run; the records inside
    a run keep their own keys."""
    from xtrack_amd import lattices
    line = lattices.collider_ring(k2f=0.03, k2d=-0.06)
    t, o, pp = build_raw_table(line.elements)
    _, key, prog, src, _ = generated_host_line(t, o, pp, line.get_length())
    rec, span, seg = key & 0x7ff, (key >> 11) & 31, key >> 16
    assert len(np.unique(rec)) == 10
    assert src.count('XT_GEN_FN void rec_') == 10
    assert 'xt::op_yoshida4(p' in src and 'xt::op_limit_rect(p' in src
    n_seg = src.count('XT_GEN_FN int32_t seg_')
    assert n_seg > 0 and set(np.unique(seg)) == set(range(n_seg + 1))
    starts = np.flatnonzero(seg)
    assert np.all(span[starts] > 1) and np.all(span[seg == 0] == 0)
    # segments do not overlap, and the same segment covers the same records
    assert np.all(np.diff(starts) >= span[starts][:-1])
    names = [type(ee).__name__ for ee in line.elements]
    for s in range(1, n_seg + 1):
        runs = {tuple(rec[e:e + span[e]]) for e in starts[seg[starts] == s]}
        assert len(runs) == 1
        for e in starts[seg[starts] == s]:
            assert not {'Bend', 'Quadrupole', 'Cavity'} & set(names[e:e + span[e]])
    # most elements of the ring are inside segments
    assert span.sum() > 0.8 * len(key)


def test_segments_losses_partial_ranges_and_monitors():
    """Runs of cheap elements tracked as segments: particles lost at every
    aperture inside the runs (at_element must name the element inside the
    segment), tracking ranges that start and stop inside segments, and
    element-by-element records, which track element by element."""
    line = segment_line()
    t, o, pp = build_raw_table(line.elements)
    _, key, _, _, _ = generated_host_line(t, o, pp, line.get_length())
    span, seg = (key >> 11) & 31, key >> 16
    pos = np.full(len(key), -1)
    for e in np.flatnonzero(seg):
        pos[e:e + span[e]] = np.arange(span[e])
    assert (span > 1).sum() >= 4 and span.sum() >= 100
    p = random_particles(800, seed=1, scale=2e-3)
    for kw in (dict(num_turns=3), dict(num_turns=2, ele_start=5, ele_stop=50),
               dict(ele_start=20, num_elements=70), dict(ele_start=16, ele_stop=27),
               dict(num_turns=2, turn_by_turn_monitor=True)):
        lg, lc, pg, pc = _run(segment_line, p, **kw)
        assert not compare_exact(pg, pc), kw
        if kw.get('turn_by_turn_monitor'):
            for ff in ('x', 'px', 'y', 'at_element', 'state'):
                assert np.array_equal(getattr(lg.record_last_track, ff),
                                      getattr(lc.record_last_track, ff), equal_nan=True)
        if kw == dict(num_turns=3):
            lost = pc.sorted_by_id()
            at = lost['at_element'][lost['state'] <= 0]
            assert len(np.unique(pos[at][pos[at] > 0])) >= 5
    lg, lc, pg, pc = _run(segment_line, p, num_turns=1, turn_by_turn_monitor='ONE_TURN_EBE')
    assert not compare_exact(pg, pc)
    for ff in ('x', 'px', 'y', 'py', 'zeta', 'at_element', 'state'):
        assert np.array_equal(getattr(lg.record_last_track, ff),
                              getattr(lc.record_last_track, ff), equal_nan=True), ff


def _hybrid_line():
    """segment_line with 80 thick quadrupoles of individual strengths
    inserted: more than 64 distinct records, so single-use records get
    structure functions (constants loaded) and the keys carry program
    offsets, while the repeated cheap records keep literal segments whose
    literal-kick fallback addresses each element's own descriptor."""
    base = segment_line(n_blocks=20, aperture=0.012)
    els = list(base.elements)
    rng = np.random.default_rng(5)
    out = []
    for k, ee in enumerate(els):
        out.append(ee)
        if k % 7 == 3 and len(out) < 10_000:
            out.append(xt.Quadrupole(length=0.3, k1=0.02 * (1 + 0.1 * rng.standard_normal()),
                                     knl=[0, 0, 0.01 * rng.standard_normal()]))
    return xt.Line(elements=out)


def test_hybrid_line_code_structures_and_segments():
    line = _hybrid_line()
    types, offsets, params = build_raw_table(line.elements)
    _, rec, _, src, entry = generated_host_line(types, offsets, params, line.get_length())
    assert len(np.unique(entry)) > 64
    assert 'struct Key { int32_t r; int64_t pc; }' in src
    assert 'xt::cptr(prog) + ent[' in src and 'seg_0' in src
    assert src.count('XT_GEN_FN void rec_') < 30
    for kw in (dict(num_turns=3), dict(ele_start=5, num_elements=40)):
        for p in (random_particles(500, seed=9, scale=5e-3), guard_edge_particles(600, seed=10)):
            _, _, pg, pc = _run(_hybrid_line, p, **kw)
            bad = compare_exact(pg, pc)
            assert not bad, bad
