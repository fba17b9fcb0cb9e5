"""The SPS lattice of the reference's test data (line JSON with its
reference particle, 8200 elements) through the JSON loader, and the kernel's
element code (host build and generated line code) against the oracle on it;
transversely shifted elements against closed forms."""
import math
import os
from collections import Counter

import numpy as np
import pytest

import xtrack_amd as xt
from helpers import compare_exact
from oracle_backend import OracleBackend, HostKernelBackend, GeneratedHostBackend
from xtrack_amd.elements import build_raw_table

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')
SPS_JSON = os.path.join(GOLDEN, 'sps_line_no_spacecharge_and_particle.json.gz')


def sps_line():
    return xt.Line.from_json(SPS_JSON)


def sps_particles(n, seed=0, n_large=None):
    """Bunch around the file's particle (x=2 mm, y=3 mm, zeta=0.2 m at
    26 GeV/c); a quarter at apertures-scale amplitudes so that every kind
    of aperture, the shifted one included, loses particles."""
    line = sps_line()
    rng = np.random.default_rng(seed)
    n_large = n // 4 if n_large is None else n_large
    x = rng.normal(2e-3, 3e-3, n)
    y = rng.normal(3e-3, 2e-3, n)
    x[:n_large] = rng.uniform(-0.1, 0.1, n_large)
    y[:n_large] = rng.uniform(-0.03, 0.03, n_large)
    return xt.Particles(p0c=line.json_particle.p0c[0], mass0=line.json_particle.mass0,
                        x=x, y=y, px=rng.normal(0, 5e-5, n), py=rng.normal(0, 5e-5, n),
                        zeta=rng.normal(0.2, 0.3, n), delta=rng.normal(0, 1e-3, n))


def test_sps_json_load():
    line = sps_line()
    assert len(line.elements) == 8200
    assert Counter(type(e).__name__ for e in line.elements) == {
        'Drift': 3537, 'DipoleEdge': 1488, 'Multipole': 1470, 'LimitRect': 1101,
        'LimitEllipse': 597, 'Cavity': 6, 'Marker': 1}
    assert line.get_length() == pytest.approx(6911.5038, abs=1e-9)
    assert line.json_particle.p0c[0] == 25983064590.009007
    # the one shifted aperture becomes a wrapped record
    types, _, _ = build_raw_table(line.elements)
    assert Counter(types.tolist())[xt.elements.T_TRANSFORMED] == 1
    i = line.element_names.index('tpst.21760_aper')
    assert types[i] == xt.elements.T_TRANSFORMED


def test_dipole_edge_linear_coefficients():
    """dipoleedge.py: with no fringe gap r21 = k tan(e1), r43 = -r21."""
    line = sps_line()
    edges = [e for e in line.elements if isinstance(e, xt.DipoleEdge)]
    for e in edges[:50]:
        assert e.hgap == 0.0 and e.fint == 0.0
        assert abs(e.r21 - e.k * math.tan(e.e1)) <= 1e-18
        assert e.r43 == -e.r21
    e = xt.DipoleEdge(k=0.01, e1=0.05, hgap=0.02, fint=0.5)
    corr = 2 * 0.01 * 0.02 * 0.5
    psi = 0.05 - corr * (1 + math.sin(0.05) ** 2) / math.cos(0.05)
    assert abs(e.r43 - (-0.01 * math.tan(psi))) <= 1e-18


@pytest.mark.parametrize('backend', [HostKernelBackend, GeneratedHostBackend],
                         ids=['interpreter', 'generated'])
def test_sps_kernel_code_vs_oracle(backend):
    p = sps_particles(400, seed=1)
    lk, lo = sps_line(), sps_line()
    lk.build_tracker(backend=backend())
    lo.build_tracker(backend=OracleBackend())
    pk, po = p.copy(), p.copy()
    lk.track(pk, num_turns=3)
    lo.track(po, num_turns=3)
    assert not compare_exact(pk, po)
    so = po.sorted_by_id()
    lost = so['state'] <= 0
    assert 20 < lost.sum() < 300
    assert len(np.unique(so['at_element'][lost])) > 5


def test_sps_losses_at_shifted_aperture():
    """Particles reaching tpst.21760_aper with x above 0.04 (its window,
    [-0.08825, 0.08825] shifted by -0.04825) are lost there, and their
    recorded x is back in the line frame."""
    line = sps_line()
    i = line.element_names.index('tpst.21760_aper')
    line.build_tracker(backend=OracleBackend())
    x = np.array([0.039, 0.0399, 0.0401, 0.041, -0.0755])
    p = xt.Particles(p0c=26e9, x=x, at_element=i)
    line.track(p, ele_start=i, num_elements=1)
    s = p.sorted_by_id()
    assert list(s['state']) == [1, 1, 0, 0, 1]
    assert np.allclose(s['x'], x, rtol=0, atol=1e-17)
    assert list(s['at_element'][2:4]) == [i, i]


def test_sps_backtrack_round_trip():
    """One turn forward then one backwards returns the surviving particles
    (DipoleEdge and the shifted aperture included)."""
    p = sps_particles(200, seed=2, n_large=0)
    line = sps_line()
    line.build_tracker(backend=OracleBackend())
    q = p.copy()
    line.track(q, num_turns=1)
    line.track(q, num_turns=1, backtrack=True)
    a, b = p.sorted_by_id(), q.sorted_by_id()
    ok = b['state'] > 0
    assert ok.sum() > 190
    # round-off of 8200 elements each way; zeta accumulates it over 6.9 km
    for ff, tol in (('x', 1e-13), ('px', 1e-15), ('y', 1e-13), ('py', 1e-15),
                    ('zeta', 1e-11), ('delta', 1e-14)):
        assert np.allclose(a[ff][ok], b[ff][ok], rtol=0, atol=tol), ff
    assert np.all(b['at_turn'][ok] == 0)


# ---------------------------------------------------------------------------
# transverse shifts (track_misalignments.h:38-110, straight frame)
def test_shifted_thin_multipole_closed_form():
    dx, dy, k1, k1s = 2e-3, -1e-3, 0.03, 0.01
    line = xt.Line(elements=[xt.Multipole(knl=[0, k1], ksl=[0, k1s], shift_x=dx, shift_y=dy)])
    line.build_tracker(backend=OracleBackend())
    rng = np.random.default_rng(3)
    x, y = rng.normal(0, 3e-3, 50), rng.normal(0, 3e-3, 50)
    p = xt.Particles(p0c=26e9, x=x, y=y)
    line.track(p)
    s = p.sorted_by_id()
    u, v = x - dx, y - dy
    # normal and skew quadrupole kick about the shifted axis
    assert np.allclose(s['px'], -(k1 * u - k1s * v), rtol=1e-14, atol=1e-19)
    assert np.allclose(s['py'], k1 * v + k1s * u, rtol=1e-14, atol=1e-19)
    assert np.allclose(s['x'], x, rtol=0, atol=1e-18)
    assert np.allclose(s['y'], y, rtol=0, atol=1e-18)


@pytest.mark.parametrize('backend', [HostKernelBackend, GeneratedHostBackend],
                         ids=['interpreter', 'generated'])
def test_shifted_elements_kernel_vs_oracle(backend):
    def mk():
        return xt.Line(elements=[
            xt.Drift(length=0.5),
            xt.Quadrupole(length=0.8, k1=0.4, shift_x=1e-3, shift_y=-2e-3),
            xt.Sextupole(length=0.3, k2=2.0, shift_y=5e-4),
            xt.Octupole(length=0.2, k3=50.0, shift_x=-3e-4),
            xt.Multipole(knl=[1e-4, 0.02], length=0.5, isthick=True, hxl=1e-4, shift_x=1e-3),
            xt.Cavity(voltage=5e6, frequency=400e6, lag=175, length=1.0, shift_x=1e-3),
            xt.DipoleEdge(k=0.01, e1=0.05, shift_y=1e-3),
            xt.LimitRect(min_x=-0.01, max_x=0.01, min_y=-0.01, max_y=0.01, shift_x=3e-3),
            xt.LimitEllipse(a=0.01, b=0.008, shift_y=-2e-3),
            xt.Drift(length=0.5)])
    rng = np.random.default_rng(4)
    n = 600
    p = xt.Particles(p0c=7e9, x=rng.normal(0, 5e-3, n), px=rng.normal(0, 1e-4, n),
                     y=rng.normal(0, 5e-3, n), py=rng.normal(0, 1e-4, n),
                     zeta=rng.normal(0, 0.05, n), delta=rng.normal(0, 1e-3, n))
    # the library backtracks on the interpreter only
    for bt in ((False, True) if backend is HostKernelBackend else (False,)):
        lk, lo = mk(), mk()
        lk.build_tracker(backend=backend())
        lo.build_tracker(backend=OracleBackend())
        pk, po = p.copy(), p.copy()
        lk.track(pk, num_turns=3, backtrack=bt)
        lo.track(po, num_turns=3, backtrack=bt)
        assert not compare_exact(pk, po)
        assert (po.state <= 0).sum() > 50
