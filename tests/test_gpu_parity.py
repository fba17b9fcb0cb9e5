"""GPU (libxtrack_amd.so) against the CPU oracle built with the same
elementary-function library: every coordinate and every loss record must be
identical (IEEE equality)."""
import numpy as np
import pytest

import xtrack_amd as xt
from helpers import mixed_line, random_particles, compare_exact
from oracle_backend import OracleBackend

pytestmark = pytest.mark.gpu


def _pair(line_fn, p):
    line_g = line_fn()
    line_g.build_tracker(device='cuda:0')
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
