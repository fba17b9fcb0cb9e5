"""Host-side checks that need no GPU: the C ABI library loads and exports
every declared symbol, the element-table builder deduplicates and rejects
what it cannot track, particle construction follows the reference formulas,
and the Line.track segmentation issues the reference's launch sequence."""
import ctypes
import math
import os
import re

import mpmath as mp
import numpy as np
import pytest

import xtrack_amd as xt
from xtrack_amd import _native, lattices
from xtrack_amd.elements import build_raw_table

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    src = open(os.path.join(ROOT, 'include', 'xtrack_amd.h')).read()
    src = re.sub(r'/\*.*?\*/', '', src, flags=re.S)
    return re.findall(r'\b(xtm_[a-z_0-9]+)\s*\(', src)


def test_library_exports_declared_symbols():
    lib = ctypes.CDLL(_native.lib_path())
    names = _declared_functions()
    assert len(names) >= 10
    for nn in names:
        assert hasattr(lib, nn), nn
    assert _native.load().xtm_abi_version() == 1


def _host_line(line):
    lib = _native.load()
    t, o, p = build_raw_table(line.elements)
    h = ctypes.c_void_p()
    st = lib.xtm_line_create(t.ctypes.data, o.ctypes.data, p.ctypes.data, len(t),
                             line.get_length(), -1, ctypes.byref(h))
    return lib, st, h


def test_builder_deduplicates_records():
    lib, st, h = _host_line(lattices.collider_ring(k2f=0.03, k2d=-0.06))
    assert st == 0
    assert lib.xtm_line_num_elements(h) == 28338
    # cavity, LimitRect, QF, QD, SF, SD, bend, two octupole signs, drift
    assert lib.xtm_line_num_records(h) == 10
    lib.xtm_line_destroy(h)


def test_host_only_line_cannot_track():
    lib, st, h = _host_line(xt.Line(elements=[xt.Drift(length=1)]))
    assert st == 0
    p = xt.Particles(x=[0.0])
    pst = _native.make_particles_struct(p)
    assert lib.xtm_track_line(h, ctypes.byref(pst), 1, 0, 1, 1, 1, 0, None, 0, None) != 0
    assert b'without a device' in lib.xtm_last_error()


@pytest.mark.parametrize('el', [
    xt.Bend(length=1.0, angle=0.1, edge_entry_model='full'),
    xt.Quadrupole(length=1.0, k1=0.1, edge_entry_active=True),
])
def test_builder_rejects_unsupported_models(el):
    lib, st, h = _host_line(xt.Line(elements=[xt.Drift(length=1), el]))
    assert st == -2
    assert b'element 1' in lib.xtm_last_error()


def test_misalignment_table_build():
    # misaligned elements are wrapped (track_local_particle_with_transformations.h)
    q = xt.Quadrupole(length=1.0, k1=0.1, shift_x=1e-3, rot_s_rad=0.2)
    types, offsets, params = build_raw_table([q])
    assert list(types) == [xt.elements.T_TRANSFORMED]
    assert list(params[:13]) == [1e-3, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.2, 1.0, 0.0, 0.0, 0.0,
                                 float(xt.elements.T_QUADRUPOLE)]
    b = xt.Bend(length=2.0, angle=0.1, rot_y_rad=1e-3, rot_shift_anchor=1.0)
    _, _, params = build_raw_table([b])
    assert list(params[3:13]) == [0.0, 1e-3, 0.0, 1.0, 0.0, 2.0, 0.1, 0.05, 1.0,
                                  float(xt.elements.T_BEND)]
    # thin multipole: no length; the anchor alone does not switch them on
    m = xt.Multipole(knl=[0, 1e-3], length=0.5, shift_y=1e-3)
    assert build_raw_table([m])[2][8] == 0.0
    assert build_raw_table([xt.Multipole(knl=[0, 1.0], rot_shift_anchor=0.3)])[0][0] == \
        xt.elements.T_MULTIPOLE
    # Drift and Marker do not allow shifts or rotations
    for ee in (xt.Drift(length=1.0, shift_x=1e-3), xt.Marker(rot_s_rad=0.1)):
        with pytest.raises(NotImplementedError):
            build_raw_table([ee])


def test_track_flags_register():
    f = xt.TrackFlags()
    f.XS_FLAG_IGNORE_LOCAL_APERTURE = True
    f.XS_FLAG_KILL_CAVITY_KICK = True
    assert int(f.get_flags_register()) == (1 << 4) | (1 << 2)


# ---------------------------------------------------------------------------
# particle construction (xtrack particles.py:1900-2030) against mpmath
def test_particles_energy_relations():
    mp.mp.prec = 200
    delta = np.array([-1e-2, 0.0, 3e-3, 0.2])
    p = xt.Particles(p0c=26e9, delta=delta, mass0=xt.PROTON_MASS_EV)
    m0 = mp.mpf(xt.PROTON_MASS_EV)
    p0c = mp.mpf(26e9)
    e0 = mp.sqrt(p0c ** 2 + m0 ** 2)
    beta0 = p0c / e0
    assert abs(p.beta0[0] - float(beta0)) < 1e-16
    assert abs(p.gamma0[0] - float(e0 / m0)) < 1e-14 * float(e0 / m0)
    for i, d in enumerate(delta):
        pc = p0c * (1 + mp.mpf(d))
        e = mp.sqrt(pc ** 2 + m0 ** 2)
        ptau = (e - e0) / p0c
        beta = pc / e
        assert abs(p.ptau[i] - float(ptau)) < 1e-15
        assert abs(p.rvv[i] - float(beta / beta0)) < 1e-15
        assert abs(p.rpp[i] - float(1 / (1 + mp.mpf(d)))) < 1e-15
    q = xt.Particles(p0c=26e9, ptau=p.ptau)
    assert np.allclose(q.delta, delta, rtol=0, atol=1e-15)
    r = xt.Particles(p0c=26e9, pzeta=p.ptau / p.beta0)
    assert np.allclose(r.delta, delta, rtol=0, atol=1e-15)


def test_particles_defaults_and_capacity():
    p = xt.Particles(x=[1e-3, 2e-3], _capacity=5)
    assert list(p.state) == [1, 1, -999999999, -999999999, -999999999]
    assert list(p.particle_id[:2]) == [0, 1]
    assert p.p0c[0] == 1e9 and p.chi[0] == 1.0 and p.charge_ratio[0] == 1.0
    with pytest.raises(NameError):
        xt.Particles(sigma=1.0)
    with pytest.raises(ValueError):
        xt.Particles(x=[1, 2], y=[1, 2, 3])


def test_particles_dict_roundtrip():
    p = xt.Particles(p0c=7e12, x=[1e-3, -1e-3], delta=[1e-4, 0], zeta=[0.1, -0.1])
    q = xt.Particles.from_dict(p.to_dict())
    for nn, vv in p.to_numpy().items():
        assert np.array_equal(vv, q.to_numpy()[nn])


def test_bend_parameter_derivation():
    b = xt.Bend(length=14.3, angle=2 * math.pi / 1104)
    assert b.h == (2 * math.pi / 1104) / 14.3 and b.k0 == 'from_h'
    assert b.raw()[1][11] == b.h
    b2 = xt.Bend(length=2.0, angle=0.1, k0=0.04)
    assert b2.raw()[1][11] == 0.04 and b2.raw()[1][19] == 0.05
    with pytest.raises(ValueError):
        xt.Bend(length=1.0, h=0.1)


def test_multipole_order_and_factorial():
    m = xt.Multipole(knl=[0, 1, 2], ksl=[0, 0, 0, 4])
    assert m.order == 3 and m.inv_factorial_order == 1 / 6
    q = xt.Quadrupole(length=1.0, k1=0.1)
    assert q.order == 5 and len(q.knl) == 6
    assert list(q.knl_rel) == [0.0]


# ---------------------------------------------------------------------------
# Line.track segmentation (tracker.py:1270-1436) with a recording backend
class _Recorder:
    space = 'host'

    def __init__(self):
        self.calls = []

    def prepare_line(self, *a):
        return None

    def track_line(self, h, p, num_turns, ele_start, num_ele_track, end_turn, reset_s,
                   flag_monitor, monitor, flags):
        self.calls.append((num_turns, ele_start, num_ele_track, bool(end_turn), flag_monitor))

    def synchronize(self):
        pass

    def monitor_context(self):
        return None


@pytest.mark.parametrize('kw,expected', [
    (dict(num_turns=3), [(1, 0, 10, True, 0), (2, 0, 10, True, 0)]),
    (dict(num_turns=1, ele_start=4), [(1, 4, 6, True, 0)]),
    (dict(num_turns=2, ele_start=4, ele_stop=7), [(1, 4, 6, True, 0), (1, 0, 7, False, 0)]),
    (dict(num_turns=1, ele_start=7, ele_stop=3), [(1, 7, 3, True, 0), (1, 0, 3, False, 0)]),
    (dict(ele_start=2, num_elements=25), [(1, 2, 8, True, 0), (1, 0, 10, True, 0),
                                          (1, 0, 7, False, 0)]),
    (dict(num_turns=2, turn_by_turn_monitor=True), [(1, 0, 10, True, 1), (1, 0, 10, True, 1)]),
])
def test_track_segmentation(kw, expected):
    line = xt.Line(elements=[xt.Drift(length=1.0) for _ in range(10)])
    rec = _Recorder()
    line.build_tracker(backend=rec)
    line.track(xt.Particles(x=[0.0, 1e-3]), **kw)
    assert rec.calls == expected


def test_track_argument_errors():
    line = xt.Line(elements=[xt.Drift(length=1.0) for _ in range(4)])
    line.build_tracker(backend=_Recorder())
    p = xt.Particles(x=[0.0])
    with pytest.raises(ValueError):
        line.track(p, num_elements=3, ele_stop=2)
    with pytest.raises(ValueError):
        line.track(p, num_elements=3, num_turns=2)
    with pytest.raises(ValueError):
        line.track(p, backtrack='bogus')
    with pytest.raises(ValueError):
        line.track(p, turn_by_turn_monitor='bogus')


def test_backtrack_sets_flag_for_the_call_only():
    """tracker.py:1222-1236: XS_FLAG_BACKTRACK (bit 0) is set while the
    request runs and restored afterwards; segmentation is unchanged."""
    line = xt.Line(elements=[xt.Drift(length=1.0) for _ in range(10)])
    rec = _Recorder()
    seen = []
    orig = rec.track_line

    def spy(h, p, num_turns, ele_start, num_ele_track, end_turn, reset_s, flag_monitor,
            monitor, flags):
        seen.append(int(flags) & 1)
        orig(h, p, num_turns, ele_start, num_ele_track, end_turn, reset_s, flag_monitor,
             monitor, flags)
    rec.track_line = spy
    line.build_tracker(backend=rec)
    line.track(xt.Particles(x=[0.0]), num_turns=2, backtrack=True)
    assert seen == [1, 1]
    assert rec.calls == [(1, 0, 10, True, 0), (1, 0, 10, True, 0)]
    line.track(xt.Particles(x=[0.0]), num_turns=1)
    assert seen[-1] == 0 and not line.track_flags.XS_FLAG_BACKTRACK


@pytest.mark.parametrize('el', [
    xt.Bend(length=1.0, angle=0.02, k1=0.01, model='rot-kick-rot', edge_entry_model='suppressed'),
    xt.Bend(length=2.0, angle=0.1, k0=0.04),
    xt.Multipole(knl=[0, 1], length=0.5, isthick=True),
    xt.Quadrupole(length=1, k1=0.1, model='drift-kick-drift-exact', shift_x=1e-3),
    xt.Cavity(voltage=1e6, frequency=4e8, lag=10, model='drift-kick-drift-exact'),
    xt.DipoleEdge(k=0.01, e1=0.1, hgap=0.02, fint=0.5),
    xt.Drift(length=1, model='exact'),
    xt.RFMultipole(knl=[1, 2], frequency=1e8),
])
def test_element_dict_round_trip(el):
    """to_dict gives constructor arguments that rebuild the same record,
    the fields behind properties (angle, k0, model, ...) included."""
    d = el.to_dict()
    cls = d.pop('__class__')
    el2 = getattr(xt, cls)(**d)
    assert np.array_equal(np.asarray(el.raw()[1]), np.asarray(el2.raw()[1]))
