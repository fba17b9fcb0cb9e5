"""Fast arithmetic (line code compiled with XTM_SPEC_FAST_ARITH) on the
host: the generated code with the re-associated polar-drift path length,
compiled with g++, against the oracle.  Without RF the fast form changes
zeta only, so every other field -- losses included -- keeps the oracle's
bits; with RF the differences stay at the rounding level."""
import numpy as np
import pytest

import xtrack_amd as xt
from helpers import mixed_line, random_particles, guard_edge_particles, compare_exact, COMPARED
from oracle_backend import OracleBackend, GeneratedHostBackend
from xtrack_amd import _native


def _pair(line_fn, p, **kw):
    lf = line_fn()
    lf.build_tracker(backend=GeneratedHostBackend(options=_native.XTM_SPEC_FAST_ARITH))
    lc = line_fn()
    lc.build_tracker(backend=OracleBackend())
    pf, pc = p.copy(), p.copy()
    lf.track(pf, **kw)
    lc.track(pc, **kw)
    return pf, pc


def _rel(a, b, ff):
    sa, sb = a.sorted_by_id(), b.sorted_by_id()
    m = (sa['state'] > 0) & (sb['state'] > 0)
    return float(np.max(np.abs(sa[ff][m] - sb[ff][m])) / np.max(np.abs(sb[ff][m])))


@pytest.mark.parametrize('maker', [lambda n, seed: random_particles(n, seed=seed, scale=4e-3),
                                   guard_edge_particles], ids=['beam', 'range-edges'])
def test_without_rf_only_zeta_changes(maker):
    """Every bend model of the mixed ring, no cavity, 6 turns: all fields
    but zeta bit-identical to the oracle (losses included); zeta within
    1e-13 of its magnitude."""
    p = maker(800, seed=21)
    pf, pc = _pair(lambda: mixed_line(n_cells=2, with_cavity=False), p, num_turns=6)
    bad = compare_exact(pf, pc, fields=[ff for ff in COMPARED if ff != 'zeta'])
    assert not bad, bad
    assert int((pc.sorted_by_id()['state'] <= 0).sum()) > 10
    assert _rel(pf, pc, 'zeta') < 1e-13


def test_with_rf_rounding_level():
    """With the cavity, zeta's rounding differences reach delta through the
    RF kick: 20 turns, identical losses, coordinates within 1e-12."""
    p = random_particles(800, seed=5, scale=2e-3)
    pf, pc = _pair(lambda: mixed_line(n_cells=2), p, num_turns=20)
    for ff in ('state', 'at_turn', 'at_element'):
        assert np.array_equal(pf.sorted_by_id()[ff], pc.sorted_by_id()[ff]), ff
    for ff in ('x', 'px', 'y', 'py', 'zeta', 'delta'):
        assert _rel(pf, pc, ff) < 1e-12, ff


def test_specialize_rejects_unknown_options():
    lib = _native.load()
    import ctypes
    line = mixed_line(n_cells=1)
    from xtrack_amd.elements import build_raw_table
    types, offsets, params = build_raw_table(line.elements)
    h = ctypes.c_void_p()
    _native.check(lib.xtm_line_create(types.ctypes.data, offsets.ctypes.data,
                                      params.ctypes.data, len(types), line.get_length(), -1,
                                      ctypes.byref(h)))
    try:
        n = ctypes.c_int64()
        assert lib.xtm_line_generate_source_ex(h, 0, 2, None, 0, ctypes.byref(n), None) != 0
        assert lib.xtm_line_specialized_options(h) == -1
        assert lib.xtm_line_generate_source_ex(h, 0, 1, None, 0, ctypes.byref(n), None) == 0
    finally:
        lib.xtm_line_destroy(h)


def test_fast_needs_compiled_code():
    from xtrack_amd.tracker import HIPBackend
    with pytest.raises(ValueError):
        HIPBackend(compile=False, arithmetic='fast')
    with pytest.raises(ValueError):
        HIPBackend(arithmetic='quick')
