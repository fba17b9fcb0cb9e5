"""The HL-LHC v1.4 lattice of the reference's test data (line JSON with its
particle; 11,843 thin elements with crab cavities as RF multipoles).  Its
224 beam-beam elements are outside this tracker's path and are replaced by
markers when the file is loaded here.  Checks: the loader, the lattice's
fractional tunes (0.31, 0.32) from the one-turn map, and the kernel's
element code (host builds) against the oracle bit for bit."""
import gzip
import json
import os
from collections import Counter

import numpy as np
import pytest

import xtrack_amd as xt
from helpers import compare_exact
from oracle_backend import OracleBackend, HostKernelBackend, GeneratedHostBackend

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')
HLLHC_JSON = os.path.join(GOLDEN, 'hllhc14_line_and_particle.json.gz')


def _load():
    with gzip.open(HLLHC_JSON, 'rt') as fh:
        return json.load(fh)


def hllhc_line():
    d = _load()
    elements = d['line']['elements']
    for nn, ee in elements.items():
        if ee['__class__'].startswith('BeamBeam'):
            elements[nn] = {'__class__': 'Marker'}
    return xt.Line.from_dict(d['line'])


def hllhc_particles(n, seed=0, n_large=None):
    """Bunch at the file's reference momentum (7 TeV protons); a fifth at
    amplitudes that reach the global aperture."""
    pd = _load()['particle']
    rng = np.random.default_rng(seed)
    n_large = n // 5 if n_large is None else n_large
    x = rng.normal(0, 5e-4, n)
    y = rng.normal(0, 5e-4, n)
    x[:n_large] = rng.uniform(-0.05, 0.05, n_large)
    return xt.Particles(p0c=pd['p0c'][0], mass0=pd['mass0'], x=x, y=y,
                        px=rng.normal(0, 2e-5, n), py=rng.normal(0, 2e-5, n),
                        zeta=rng.normal(0, 0.08, n), delta=rng.normal(0, 1e-4, n))


def test_hllhc_json_load():
    line = hllhc_line()
    assert Counter(type(e).__name__ for e in line.elements) == {
        'Drift': 6016, 'Multipole': 5579, 'Marker': 224, 'RFMultipole': 16, 'Cavity': 8}
    assert line.get_length() == pytest.approx(26658.8832, abs=1e-6)
    crab = [e for e in line.elements if isinstance(e, xt.RFMultipole)][0]
    # order 0 in the file becomes the default order 5 (_common.py:438)
    assert crab.order == 5 and crab.frequency == pytest.approx(400789602.58620286)


def test_hllhc_tunes():
    """Fractional tunes of the bare lattice from the oracle's one-turn
    map: the HL-LHC collision optics' 0.31 / 0.32."""
    from xtrack_amd import optics
    line = hllhc_line()
    line.build_tracker(backend=OracleBackend())
    pd = _load()['particle']
    tw = optics.twiss(line, xt.Particles(p0c=pd['p0c'][0], mass0=pd['mass0']))
    assert abs(tw['qx'] - 0.31) < 2e-5
    assert abs(tw['qy'] - 0.32) < 2e-5


@pytest.mark.parametrize('backend', [HostKernelBackend, GeneratedHostBackend],
                         ids=['interpreter', 'generated'])
def test_hllhc_kernel_code_vs_oracle(backend):
    p = hllhc_particles(200, seed=1)
    lk, lo = hllhc_line(), hllhc_line()
    lk.build_tracker(backend=backend())
    lo.build_tracker(backend=OracleBackend())
    pk, po = p.copy(), p.copy()
    lk.track(pk, num_turns=2)
    lo.track(po, num_turns=2)
    assert not compare_exact(pk, po)
    assert (po.state <= 0).sum() > 5
