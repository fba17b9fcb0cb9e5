"""Fast arithmetic on the GPU: the compiled line code built with
XTM_SPEC_FAST_ARITH against the same generated code compiled for the host
(identical bits: the device's in-range quotient sequences return the IEEE
results the host computes), and against the oracle without RF (every field
but zeta identical)."""
import numpy as np
import pytest

from helpers import mixed_line, random_particles, guard_edge_particles, compare_exact, COMPARED
from oracle_backend import OracleBackend, GeneratedHostBackend
from xtrack_amd import _native

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('case', ['beam', 'range-edges'])
def test_fast_device_equals_host_build(case):
    print(f'[fast arithmetic] {case}', flush=True)
    p = (random_particles(2000, seed=31, scale=4e-3) if case == 'beam'
         else guard_edge_particles(3000, seed=32))
    lg = mixed_line(n_cells=3, aperture=0.02 if case == 'beam' else 0.95)
    lg.build_tracker(device='cuda:0', arithmetic='fast')
    lib = lg.tracker.backend.lib
    assert lib.xtm_line_specialized_options(lg.tracker._handle.ptr) == _native.XTM_SPEC_FAST_ARITH
    lh = mixed_line(n_cells=3, aperture=0.02 if case == 'beam' else 0.95)
    lh.build_tracker(backend=GeneratedHostBackend(options=_native.XTM_SPEC_FAST_ARITH))
    pg, ph = p.copy(_context='cuda:0'), p.copy()
    lg.track(pg, num_turns=8)
    lh.track(ph, num_turns=8)
    bad = compare_exact(pg, ph)
    assert not bad, bad


def test_fast_without_rf_only_zeta_differs_from_oracle():
    print('[fast arithmetic] no RF vs oracle', flush=True)
    p = random_particles(3000, seed=33, scale=4e-3)
    lg = mixed_line(n_cells=3, with_cavity=False)
    lg.build_tracker(device='cuda:0', arithmetic='fast')
    lc = mixed_line(n_cells=3, with_cavity=False)
    lc.build_tracker(backend=OracleBackend())
    pg, pc = p.copy(_context='cuda:0'), p.copy()
    lg.track(pg, num_turns=10)
    lc.track(pc, num_turns=10)
    bad = compare_exact(pg, pc, fields=[ff for ff in COMPARED if ff != 'zeta'])
    assert not bad, bad
    g, c = pg.sorted_by_id(), pc.sorted_by_id()
    m = c['state'] > 0
    assert int((~m).sum()) > 20
    err = np.max(np.abs(g['zeta'][m] - c['zeta'][m])) / np.max(np.abs(c['zeta'][m]))
    print(f'zeta rel err {err:.2e}', flush=True)
    assert err < 1e-13
