"""Beam elements on the tracking path, with the constructor arguments and
parameter derivation of the reference classes:

  Drift         xtrack/beam_elements/drift.py:14-60
  Multipole     multipole.py:27-175
  Quadrupole    quadrupole.py:25-160
  Bend          bend.py:24-100, _common.py:559-743
  Sextupole     sextupole.py, Octupole octupole.py (same pattern as Quadrupole)
  Cavity        cavity.py:21-171
  LimitRect     limit_rect.py, LimitEllipse limit_ellipse.py:12-80
  DipoleEdge    dipole_edge.py:12-150
  RFMultipole   rf_multipole.py:11-147, _common.py:419-541
  Marker        marker.py

Each element serialises itself into the raw element table read by the HIP
library (layout in include/xtrack_amd.h).  Misalignment fields exist on
every element as in the reference (base_element.py:274-282); an element
with a non-zero one is wrapped in a TRANSFORMED record, tracked with the
frame transformations of track_misalignments.h.  Drift and Marker do not
allow them (allow_rot_and_shift = False in the reference).
"""
import math

import numpy as np

from .constants import UNLIMITED

# element type ids (include/xtrack_amd.h)
T_MARKER, T_DRIFT, T_MULTIPOLE, T_QUADRUPOLE, T_BEND = 0, 1, 2, 3, 4
T_SEXTUPOLE, T_OCTUPOLE, T_CAVITY, T_LIMITRECT, T_LIMITELLIPSE = 5, 6, 7, 8, 9
T_DIPOLEEDGE, T_TRANSFORMED, T_RFMULTIPOLE, T_SLICE = 10, 11, 12, 13

DEFAULT_MULTIPOLE_ORDER = 5          # _common.py:18

# model / integrator names <-> indices (_common.py:20-74)
_MODEL_DRIFT = {'adaptive': 0, 'expanded': 1, 'exact': 2}
_MODEL_CURVED = {'adaptive': 0, 'full': 1, 'bend-kick-bend': 2, 'rot-kick-rot': 3,
                 'mat-kick-mat': 4, 'drift-kick-drift-exact': 5,
                 'drift-kick-drift-expanded': 6, 'rot-kick-rot-low-order': 7,
                 'rot-kick-rot-high-order': 8, 'expanded': 4}
_MODEL_STRAIGHT = {k: v for k, v in _MODEL_CURVED.items() if v not in (2, 3)}
_MODEL_RF = {k: v for k, v in _MODEL_STRAIGHT.items() if v not in (1, 4)}
_INTEGRATOR = {'adaptive': 0, 'teapot': 1, 'yoshida4': 2, 'uniform': 3}
_EDGE_MODEL = {'suppressed': -1, 'linear': 0, 'full': 1, 'dipole-only': 2}

_MISALIGN = ('shift_x', 'shift_y', 'shift_s', 'rot_x_rad', 'rot_y_rad', 'rot_s_rad',
             'rot_s_rad_no_frame', 'rot_shift_anchor')


def _name_of(table, value):
    for k, v in table.items():
        if v == value and k != 'expanded':
            return k
    return value


def _lookup(table, value, what):
    if isinstance(value, (int, np.integer)) and not isinstance(value, bool):
        if int(value) in table.values():
            return int(value)
    try:
        return table[value]
    except KeyError:
        raise ValueError(f'Invalid {what}: {value}')


class BeamElement:
    """Base class: holds fields in ``self.__dict__`` and provides
    ``to_dict``/``copy``.  Subclasses set ``_type`` and ``isthick``."""
    _type = None
    isthick = False
    _static_thick = False   # class declares isthick = True (tracker.py:681)

    def _init_misalignment(self, kwargs):
        for nn in _MISALIGN:
            setattr(self, nn, float(kwargs.pop(nn, 0.0)))

    # track_local_particle_with_transformations.h:170-204: the frame
    # transformations run when any of these is non-zero (the anchor alone
    # does not switch them on)
    _ACTIVE_MISALIGN = ('shift_x', 'shift_y', 'shift_s', 'rot_x_rad', 'rot_y_rad', 'rot_s_rad',
                        'rot_s_rad_no_frame')

    def _transform_prefix(self):
        """None for an aligned element, else the TRANSFORMED record prefix:
        the misalignment fields, the element's length (IS_THICK /
        IS_THICK_DYNAMIC classes), angle and h, and whether the class
        curves the reference frame (base_element.py:83-125)."""
        if not any(getattr(self, nn) != 0.0 for nn in self._ACTIVE_MISALIGN):
            return None
        if self._type in (T_DRIFT, T_MARKER):
            raise NotImplementedError(
                f'{type(self).__name__} does not allow shifts or rotations')
        if self._type == T_MULTIPOLE:
            length = self.length if self.isthick else 0.0
        elif self.isthick:
            length = self.length
        else:
            length = 0.0
        curved = self._type == T_BEND
        angle = self.angle if curved else 0.0
        h = self.h if curved else 0.0
        return [self.shift_x, self.shift_y, self.shift_s, self.rot_x_rad, self.rot_y_rad,
                self.rot_s_rad_no_frame, self.rot_shift_anchor, self.rot_s_rad,
                float(length), float(angle), float(h), 1.0 if curved else 0.0,
                float(self._type)]

    def copy(self):
        out = type(self).__new__(type(self))
        out.__dict__.update({k: (v.copy() if isinstance(v, np.ndarray) else v)
                             for k, v in self.__dict__.items()})
        return out

    # derived quantities that the constructors reject as arguments
    _NOT_SERIALISED = ('h',)

    def to_dict(self):
        """Constructor arguments that rebuild the element: public fields,
        and the fields held behind properties (angle, k0, model,
        integrator, edge models, isthick, ...) under the property name."""
        out = {'__class__': type(self).__name__}
        for k, v in self.__dict__.items():
            if k.startswith('_'):
                name = k[1:]
                if (name in self._NOT_SERIALISED or
                        not isinstance(getattr(type(self), name, None), property)):
                    continue
                k, v = name, getattr(self, name)
            out[k] = v.tolist() if isinstance(v, np.ndarray) else v
        return out

    def raw(self):
        """(type id, list of float parameters) for the raw element table."""
        raise NotImplementedError

    def __repr__(self):
        fields = ', '.join(f'{k}={v!r}' for k, v in self.to_dict().items() if k != '__class__')
        return f'{type(self).__name__}({fields})'


class Marker(BeamElement):
    _type = T_MARKER

    def __init__(self, **kwargs):
        self._init_misalignment(kwargs)
        if kwargs:
            raise NameError(f'Invalid argument(s): {set(kwargs)}')

    def raw(self):
        return T_MARKER, []


class Drift(BeamElement):
    _type = T_DRIFT
    isthick = True
    _static_thick = True

    def __init__(self, length=None, model=None, **kwargs):
        self._init_misalignment(kwargs)
        if kwargs:
            raise NameError(f'Invalid argument(s): {set(kwargs)}')
        self.length = float(length) if length else 0.0
        self._model = 0
        if model is not None:
            self.model = model

    @property
    def model(self):
        return _name_of(_MODEL_DRIFT, self._model)

    @model.setter
    def model(self, value):
        self._model = _lookup(_MODEL_DRIFT, value, 'model')

    def raw(self):
        return T_DRIFT, [self.length, float(self._model)]


def _prepare_multipolar(order, knl, ksl):
    """_common.py:476-537: pad knl/ksl to a common length >= order+1."""
    order = order or 0
    lengths = [len(k) if k is not None else 0 for k in (knl, ksl)]
    target_len = max(order + 1, *lengths)
    out = []
    for k in (knl, ksl):
        arr = np.zeros(target_len, dtype=np.float64)
        if k is not None:
            arr[:len(k)] = np.asarray(k, dtype=np.float64)
        out.append(arr)
    order = target_len - 1
    return order, out[0], out[1], 1.0 / math.factorial(order)


def _prepare_rel(knl_rel, ksl_rel):
    """_common.py:545-556: default [0]; pad to a common length."""
    knl_rel = [0.0] if knl_rel is None else list(knl_rel)
    ksl_rel = [0.0] if ksl_rel is None else list(ksl_rel)
    n = max(len(knl_rel), len(ksl_rel))
    if n == 0:
        raise ValueError('knl_rel/ksl_rel must have at least one entry')
    knl_rel = knl_rel + [0.0] * (n - len(knl_rel))
    ksl_rel = ksl_rel + [0.0] * (n - len(ksl_rel))
    return np.asarray(knl_rel, dtype=np.float64), np.asarray(ksl_rel, dtype=np.float64)


class _Magnet(BeamElement):
    """Common storage of the magnet classes; serialises the shared raw
    layout XTM_M_* (include/xtrack_amd.h)."""
    _model_table = _MODEL_STRAIGHT
    _default_order = DEFAULT_MULTIPOLE_ORDER

    def _init_magnet(self, kwargs, main_fields=()):
        self._init_misalignment(kwargs)
        order = kwargs.pop('order', None)
        knl = kwargs.pop('knl', None)
        ksl = kwargs.pop('ksl', None)
        kwargs.pop('inv_factorial_order', None)   # derived from order (_common.py:538-539)
        if self._default_order is not None:
            order = order or self._default_order
        self._order, self.knl, self.ksl, self.inv_factorial_order = \
            _prepare_multipolar(order, knl, ksl)
        self.knl_rel, self.ksl_rel = _prepare_rel(kwargs.pop('knl_rel', None),
                                                  kwargs.pop('ksl_rel', None))
        self.main_is_skew = bool(kwargs.pop('main_is_skew', False))
        self.length = float(kwargs.pop('length', 0.0) or 0.0)
        self.num_multipole_kicks = int(kwargs.pop('num_multipole_kicks', 0))
        self.radiation_flag = int(kwargs.pop('radiation_flag', 0))
        self.delta_taper = float(kwargs.pop('delta_taper', 0.0))
        for nn in main_fields:
            setattr(self, nn, float(kwargs.pop(nn, 0.0)))
        self._model = 0
        self._integrator = 0
        model = kwargs.pop('model', None)
        integrator = kwargs.pop('integrator', None)
        return model, integrator

    def _finish_init(self, model, integrator, kwargs):
        if model is not None:
            self.model = model
        if integrator is not None:
            self.integrator = integrator
        if kwargs:
            raise NameError(f'Invalid argument(s): {set(kwargs)}')

    @property
    def order(self):
        return self._order

    @property
    def model(self):
        return _name_of(self._model_table, self._model)

    @model.setter
    def model(self, value):
        self._model = _lookup(self._model_table, value, 'model')

    @property
    def integrator(self):
        return _name_of(_INTEGRATOR, self._integrator)

    @integrator.setter
    def integrator(self, value):
        self._integrator = _lookup(_INTEGRATOR, value, 'integrator')

    def _raw_fields(self):
        return {}

    def raw(self):
        if self.radiation_flag:
            raise NotImplementedError('radiation is not supported by the HIP tracker')
        hdr = [0.0] * 34
        hdr[0] = self.length
        hdr[1] = float(self._order)
        hdr[2] = self.inv_factorial_order
        hdr[3] = float(len(self.knl_rel))
        hdr[4] = float(self._model)
        hdr[5] = float(self._integrator)
        hdr[6] = float(self.num_multipole_kicks)
        hdr[7] = float(self.main_is_skew)
        hdr[32] = self.delta_taper
        hdr[33] = float(self.radiation_flag)
        for idx, val in self._raw_fields().items():
            hdr[idx] = float(val)
        arrays = np.concatenate([self.knl, self.ksl, self.knl_rel, self.ksl_rel])
        return self._type, hdr + arrays.tolist()


class Multipole(_Magnet):
    _type = T_MULTIPOLE
    _default_order = None

    def __init__(self, **kwargs):
        if 'bal' in kwargs:
            raise ValueError('`bal` not supported anymore')
        if kwargs.get('hyl', 0.0) != 0.0:
            raise ValueError('hyl is not supported anymore')
        kwargs.pop('hyl', None)
        isthick = kwargs.pop('isthick', False)
        self.hxl = float(kwargs.pop('hxl', 0.0))
        self.main_order = int(kwargs.pop('main_order', 0))
        model, integrator = self._init_magnet(kwargs)
        self._isthick = int(bool(isthick))
        self._finish_init(model, integrator, kwargs)

    @property
    def isthick(self):
        return bool(self._isthick > 0)

    @isthick.setter
    def isthick(self, value):
        self._isthick = int(bool(value))

    def _raw_fields(self):
        return {8: self.main_order, 9: self._isthick, 10: self.hxl}


class Quadrupole(_Magnet):
    _type = T_QUADRUPOLE
    isthick = True
    _static_thick = True

    def __init__(self, **kwargs):
        self.edge_entry_active = int(bool(kwargs.pop('edge_entry_active', False)))
        self.edge_exit_active = int(bool(kwargs.pop('edge_exit_active', False)))
        model, integrator = self._init_magnet(kwargs, ('k1', 'k1s'))
        self._finish_init(model, integrator, kwargs)

    def _raw_fields(self):
        return {12: self.k1, 16: self.k1s, 20: self.edge_entry_active,
                21: self.edge_exit_active}


class Sextupole(_Magnet):
    _type = T_SEXTUPOLE
    isthick = True
    _static_thick = True

    def __init__(self, **kwargs):
        self.edge_entry_active = int(bool(kwargs.pop('edge_entry_active', False)))
        self.edge_exit_active = int(bool(kwargs.pop('edge_exit_active', False)))
        model, integrator = self._init_magnet(kwargs, ('k2', 'k2s'))
        self._finish_init(model, integrator, kwargs)

    def _raw_fields(self):
        return {13: self.k2, 17: self.k2s, 20: self.edge_entry_active,
                21: self.edge_exit_active}


class Octupole(_Magnet):
    _type = T_OCTUPOLE
    isthick = True
    _static_thick = True

    def __init__(self, **kwargs):
        self.edge_entry_active = int(bool(kwargs.pop('edge_entry_active', False)))
        self.edge_exit_active = int(bool(kwargs.pop('edge_exit_active', False)))
        model, integrator = self._init_magnet(kwargs, ('k3', 'k3s'))
        self._finish_init(model, integrator, kwargs)

    def _raw_fields(self):
        return {14: self.k3, 18: self.k3s, 20: self.edge_entry_active,
                21: self.edge_exit_active}


class Bend(_Magnet):
    """Sector bend; ``h = angle/length`` and ``k0 = h`` unless ``k0`` is
    given (_common.py:601-682)."""
    _type = T_BEND
    isthick = True
    _static_thick = True
    _model_table = _MODEL_CURVED

    def __init__(self, **kwargs):
        if 'h' in kwargs:
            raise ValueError('Setting `h` directly is not allowed. Set `length` and `angle` instead.')
        angle = float(kwargs.pop('angle', 0.0))
        k0 = kwargs.pop('k0', 'from_h')
        self.edge_entry_active = int(kwargs.pop('edge_entry_active', 1))
        self.edge_exit_active = int(kwargs.pop('edge_exit_active', 1))
        self._edge_entry_model = _lookup(_EDGE_MODEL, kwargs.pop('edge_entry_model', 'linear'),
                                         'edge model')
        self._edge_exit_model = _lookup(_EDGE_MODEL, kwargs.pop('edge_exit_model', 'linear'),
                                        'edge model')
        for nn in ('edge_entry_angle', 'edge_exit_angle', 'edge_entry_angle_fdown',
                   'edge_exit_angle_fdown', 'edge_entry_fint', 'edge_exit_fint',
                   'edge_entry_hgap', 'edge_exit_hgap'):
            setattr(self, nn, float(kwargs.pop(nn, 0.0)))
        model, integrator = self._init_magnet(kwargs, ('k1', 'k2'))
        self._angle = angle
        self._k0_from_h = True
        self._k0 = 0.0
        self._update_h()
        self.k0 = k0
        self._finish_init(model, integrator, kwargs)

    def _update_h(self):
        if self.length != 0:
            self._h = self._angle / self.length
        else:
            self._h = 0.0
        if self._k0_from_h:
            self._k0 = self._h

    @property
    def angle(self):
        return self._angle

    @angle.setter
    def angle(self, value):
        self._angle = float(value)
        self._update_h()

    @property
    def h(self):
        return self._h

    @property
    def k0(self):
        return 'from_h' if self._k0_from_h else self._k0

    @k0.setter
    def k0(self, value):
        if isinstance(value, str):
            if value != 'from_h':
                raise ValueError("k0 can only be set to 'from_h' as a string")
            self._k0_from_h = True
            self._k0 = self._h
        else:
            self._k0_from_h = False
            self._k0 = float(value)

    @property
    def edge_entry_model(self):
        return _name_of(_EDGE_MODEL, self._edge_entry_model)

    @property
    def edge_exit_model(self):
        return _name_of(_EDGE_MODEL, self._edge_exit_model)

    def _raw_fields(self):
        return {11: self._k0, 12: self.k1, 13: self.k2, 19: self._h,
                20: self.edge_entry_active, 21: self.edge_exit_active,
                22: self._edge_entry_model, 23: self._edge_exit_model,
                24: self.edge_entry_angle, 25: self.edge_exit_angle,
                26: self.edge_entry_angle_fdown, 27: self.edge_exit_angle_fdown,
                28: self.edge_entry_fint, 29: self.edge_exit_fint,
                30: self.edge_entry_hgap, 31: self.edge_exit_hgap}


class Cavity(BeamElement):
    _type = T_CAVITY
    isthick = True
    _static_thick = True

    def __init__(self, **kwargs):
        self._init_misalignment(kwargs)
        self.length = float(kwargs.pop('length', 0.0))
        self.voltage = float(kwargs.pop('voltage', 0.0))
        self.frequency = float(kwargs.pop('frequency', 0.0))
        self.harmonic = float(kwargs.pop('harmonic', 0.0))
        if self.frequency != 0 and self.harmonic != 0:
            raise ValueError('Cannot set both frequency and harmonic.')
        self.lag = float(kwargs.pop('lag', 0.0))
        self.phase = float(kwargs.pop('phase', 0.0))
        self.lag_taper = float(kwargs.pop('lag_taper', 0.0))
        self.phase_taper = float(kwargs.pop('phase_taper', 0.0))
        self.absolute_time = int(kwargs.pop('absolute_time', 0))
        self.num_kicks = int(kwargs.pop('num_kicks', 0))
        self._model = _lookup(_MODEL_RF, kwargs.pop('model', 'adaptive'), 'model')
        self._integrator = _lookup(_INTEGRATOR, kwargs.pop('integrator', 'adaptive'),
                                   'integrator')
        if kwargs:
            raise NameError(f'Invalid argument(s): {set(kwargs)}')

    @property
    def model(self):
        return _name_of(_MODEL_RF, self._model)

    @property
    def integrator(self):
        return _name_of(_INTEGRATOR, self._integrator)

    def raw(self):
        return T_CAVITY, [self.length, self.voltage, self.frequency, self.harmonic,
                          self.lag, self.phase, float(self.absolute_time),
                          float(self.num_kicks), float(self._model),
                          float(self._integrator), self.lag_taper, self.phase_taper]


class LimitRect(BeamElement):
    _type = T_LIMITRECT

    def __init__(self, min_x=-UNLIMITED, max_x=UNLIMITED, min_y=-UNLIMITED,
                 max_y=UNLIMITED, **kwargs):
        self._init_misalignment(kwargs)
        if kwargs:
            raise NameError(f'Invalid argument(s): {set(kwargs)}')
        self.min_x, self.max_x = float(min_x), float(max_x)
        self.min_y, self.max_y = float(min_y), float(max_y)

    def raw(self):
        return T_LIMITRECT, [self.min_x, self.max_x, self.min_y, self.max_y]


class LimitEllipse(BeamElement):
    _type = T_LIMITELLIPSE

    def __init__(self, a=None, b=None, a_squ=None, b_squ=None, **kwargs):
        self._init_misalignment(kwargs)
        a_b_squ = kwargs.pop('a_b_squ', None)
        if kwargs:
            raise NameError(f'Invalid argument(s): {set(kwargs)}')
        if a is None and a_squ is None:
            a = UNLIMITED
        if b is None and b_squ is None:
            b = UNLIMITED
        if a is not None:
            a_squ = a * a
        if b is not None:
            b_squ = b * b
        if a_b_squ is None:
            a_b_squ = a_squ * b_squ
        if not (a_squ > 0.0 and b_squ > 0.0):
            raise ValueError('a_squ and b_squ have to be positive definite')
        self.a_squ, self.b_squ, self.a_b_squ = float(a_squ), float(b_squ), float(a_b_squ)

    @property
    def a(self):
        return math.sqrt(self.a_squ)

    @property
    def b(self):
        return math.sqrt(self.b_squ)

    def raw(self):
        return T_LIMITELLIPSE, [self.a_squ, self.b_squ, self.a_b_squ]


class DipoleEdge(BeamElement):
    """Thin dipole edge (dipole_edge.py).  The linear model's kick
    coefficients r21, r43 are derived from k, e1, e1_fd, hgap and fint
    exactly as the reference derives them (numpy, dipole_edge.py:123-131)."""
    _type = T_DIPOLEEDGE
    _MODELS = {'linear': 0, 'full': 1, 'suppressed': -1}
    _SIDES = {'entry': 0, 'exit': 1}

    def __init__(self, k=None, e1=None, e1_fd=None, hgap=None, fint=None, model=None,
                 side=None, **kwargs):
        self._init_misalignment(kwargs)
        if 'h' in kwargs:                       # backward-compatible name
            if k is not None:
                raise ValueError('k and h given')
            k = kwargs.pop('h')
        self.delta_taper = float(kwargs.pop('delta_taper', 0.0))
        for nn in ('r21', 'r43'):               # derived; recomputed below
            kwargs.pop(nn, None)
        if kwargs:
            raise NameError(f'Invalid argument(s): {set(kwargs)}')
        self._k = float(k) if k is not None else 0.0
        self._e1 = float(e1) if e1 is not None else 0.0
        self._e1_fd = float(e1_fd) if e1_fd is not None else 0.0
        self._hgap = float(hgap) if hgap is not None else 0.0
        self._fint = float(fint) if fint is not None else 0.0
        self._model = self._MODELS['linear'] if model is None else _lookup(self._MODELS, model, 'model')
        self._side = 0 if side is None else _lookup(self._SIDES, side, 'side')
        self._update_r21_r43()

    def _update_r21_r43(self):
        k, e1, hgap, fint = (np.float64(v) for v in (self._k, self._e1, self._hgap, self._fint))
        corr = np.float64(2.0) * k * hgap * fint
        r21 = k * np.tan(e1)
        e1_v = e1 + np.float64(self._e1_fd)
        temp = corr / np.cos(e1_v) * (np.float64(1) + np.sin(e1_v) * np.sin(e1_v))
        r43 = -k * np.tan(e1_v - temp)
        self.r21, self.r43 = float(r21), float(r43)

    def _prop(name):
        def get(self):
            return getattr(self, '_' + name)

        def set_(self, value):
            setattr(self, '_' + name, float(value))
            self._update_r21_r43()
        return property(get, set_)

    k = _prop('k')
    e1 = _prop('e1')
    e1_fd = _prop('e1_fd')
    hgap = _prop('hgap')
    fint = _prop('fint')
    del _prop

    @property
    def model(self):
        return _name_of(self._MODELS, self._model)

    @property
    def side(self):
        return _name_of(self._SIDES, self._side)

    def to_dict(self):
        return {'__class__': 'DipoleEdge', 'k': self._k, 'e1': self._e1, 'e1_fd': self._e1_fd,
                'hgap': self._hgap, 'fint': self._fint, 'model': self.model,
                'side': self.side, 'delta_taper': self.delta_taper}

    def raw(self):
        return T_DIPOLEEDGE, [self.r21, self.r43, float(self._model), self.delta_taper]


class RFMultipole(BeamElement):
    """Thin multipole whose strengths and energy kick are modulated with the
    arrival time (rf_multipole.py:11-147).  The arrays knl, ksl, pn, ps
    (degrees, deprecated), phase_n, phase_s share one length, at least
    order + 1 with order defaulting to 5 (_common.py:419-541)."""
    _type = T_RFMULTIPOLE
    _ARRAYS = ('knl', 'ksl', 'pn', 'ps', 'phase_n', 'phase_s')

    def __init__(self, **kwargs):
        self._init_misalignment(kwargs)
        self.voltage = float(kwargs.pop('voltage', 0.0))
        self.frequency = float(kwargs.pop('frequency', 0.0))
        self.lag = float(kwargs.pop('lag', 0.0))
        self.phase = float(kwargs.pop('phase', 0.0))
        self.absolute_time = int(kwargs.pop('absolute_time', 0))
        order = kwargs.pop('order', None)
        kwargs.pop('inv_factorial_order', None)
        arrays = {nn: kwargs.pop(nn, None) for nn in self._ARRAYS}
        if kwargs:
            raise NameError(f'Invalid argument(s): {set(kwargs)}')
        order = order or DEFAULT_MULTIPOLE_ORDER
        lengths = [len(v) if v is not None else 0 for v in arrays.values()]
        target_len = max(order + 1, *lengths)
        for nn, vv in arrays.items():
            arr = np.zeros(target_len, dtype=np.float64)
            if vv is not None:
                arr[:len(vv)] = np.asarray(vv, dtype=np.float64)
            setattr(self, nn, arr)
        self._order = target_len - 1
        self.inv_factorial_order = 1.0 / math.factorial(self._order)

    @property
    def order(self):
        return self._order

    def raw(self):
        head = [self.voltage, self.frequency, self.lag, self.phase, float(self._order)]
        return T_RFMULTIPOLE, head + [float(v) for nn in self._ARRAYS
                                      for v in getattr(self, nn)]


ELEMENT_CLASSES = {c.__name__: c for c in (Marker, Drift, Multipole, Quadrupole, Bend,
                                           Sextupole, Octupole, Cavity, LimitRect,
                                           LimitEllipse, DipoleEdge, RFMultipole)}

# keys of serialised elements that carry no tracking parameter (naming and
# aperture bookkeeping of the reference's environment)
_METADATA_KEYS = ('prototype', 'name_associated_aperture', 'parent_name', 'extra')


def build_raw_table(elements):
    """Concatenate element records into (types int32, offsets int64,
    params float64) arrays for the C ABI."""
    types = np.empty(len(elements), dtype=np.int32)
    offsets = np.empty(len(elements), dtype=np.int64)
    chunks = []
    pos = 0
    cache = {}
    for ii, ee in enumerate(elements):
        key = id(ee)
        if key not in cache:
            prefix = ee._transform_prefix()
            tt, pp = ee.raw()
            pp = np.asarray(pp, dtype=np.float64)
            if prefix is not None:
                pp = np.concatenate([prefix, pp])
                tt = T_TRANSFORMED
            cache[key] = (tt, pp)
        tt, pp = cache[key]
        types[ii] = tt
        offsets[ii] = pos
        chunks.append(pp)
        pos += len(pp)
    params = np.concatenate(chunks) if chunks else np.zeros(1)
    if len(params) == 0:
        params = np.zeros(1)
    return types, offsets, params
