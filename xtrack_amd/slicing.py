"""Slicing of thick elements (xtrack/slicing.py:22-470) and the slice
element classes (beam_elements/slice_elements_{thin,thick,drift,edge}.py).

``Line.slice_thick_elements(slicing_strategies)`` replaces each thick
element chosen by the strategies with, in order: an ``<name>_entry``
marker, the entry edge slice (magnets with edges), drift and thin slices
(thin mode) or thick slices (thick mode) with the scheme's weights, the
exit edge slice and an ``<name>_exit`` marker, named as the reference
names them.  Each slice refers to its parent element, which stays in the
line's element dictionary.  Aperture replicas of sliced elements are not
made (the apertures of the tracked path are separate elements here).
"""
import re
from itertools import zip_longest

from . import elements as E


class ElementSlicingScheme:
    def __init__(self, slicing_order, mode='thin'):
        if slicing_order < 1:
            raise ValueError('A slicing scheme must have at least one slice.')
        if mode not in ('thin', 'thick'):
            raise ValueError(f'Unknown slicing mode: {mode}')
        self.mode = mode
        # thick mode counts the thick slices, i.e. the "drifts"
        self.slicing_order = slicing_order - 1 if mode == 'thick' else slicing_order

    def iter_weights(self, element_length=None):
        """(weight, is_drift) in order: drift, element, drift, ...; thick
        mode yields the drift weights only (slicing.py:52-78)."""
        for dw, ew in zip_longest(self.drift_weights(element_length),
                                  self.element_weights(element_length), fillvalue=None):
            yield dw, True
            if ew is None:
                break
            if self.mode == 'thin':
                yield ew, False

    def __repr__(self):
        return f'{type(self).__name__}({self.slicing_order})'


class Uniform(ElementSlicingScheme):
    """Equal element slices; equal drifts (slicing.py:81-92)."""

    def element_weights(self, element_length=None):
        if self.slicing_order == 0 and self.mode == 'thick':
            return [1.]
        return [1. / self.slicing_order] * self.slicing_order

    def drift_weights(self, element_length=None):
        if self.slicing_order == 0:
            return [1.]
        n = self.slicing_order + 1
        return [1. / n] * n


class Teapot(ElementSlicingScheme):
    """TEAPOT drift weights: 1/(2(n+1)) at the ends, n/(n^2-1) inside
    (slicing.py:95-114)."""

    def element_weights(self, element_length=None):
        if self.slicing_order == 0 and self.mode == 'thick':
            return [1.]
        return [1. / self.slicing_order] * self.slicing_order

    def drift_weights(self, element_length=None):
        if self.slicing_order == 0:
            return [1.]
        if self.slicing_order == 1:
            return [0.5, 0.5]
        edge = 1. / (2 * (1 + self.slicing_order))
        middle = self.slicing_order / (self.slicing_order ** 2 - 1)
        return [edge] + [middle] * (self.slicing_order - 1) + [edge]


class Strategy:
    """Which elements a scheme applies to, by name (regular expression
    unless exact) and / or type (slicing.py:160-212).  Drifts match only
    when named or typed explicitly."""

    def __init__(self, slicing, name=None, element_type=None, exact=False):
        if name is None and element_type is None:
            exact = True
        self.regex = not exact
        self.match_name = re.compile(name) if (name is not None and self.regex) else name
        self.element_type = element_type
        self.slicing = slicing

    def _match_on_name(self, name):
        if self.regex:
            return self.match_name.match(name)
        return self.match_name == name

    def match_element(self, name, element):
        if isinstance(element, E.Drift):
            if self.match_name and not self.element_type:
                return bool(self._match_on_name(name))
            if self.element_type and not self.match_name:
                return isinstance(element, self.element_type)
            if self.match_name and self.element_type:
                return bool(self._match_on_name(name)) and isinstance(element,
                                                                       self.element_type)
            return False
        matched = True
        if self.match_name:
            matched = matched and bool(self._match_on_name(name))
        if self.element_type:
            matched = matched and isinstance(element, self.element_type)
        return matched


# ---------------------------------------------------------------------------
# slice elements

K_THIN, K_THICK, K_DRIFT, K_ENTRY, K_EXIT = 0, 1, 2, 3, 4


class _SliceBase(E.BeamElement):
    """A slice of `_parent` with a weight and its offset from the parent's
    entry (slice_base.py:22-60)."""
    _type = E.T_SLICE
    _kind = None
    _parent_class = None
    isthick = False

    def __init__(self, _parent=None, weight=0.0, slice_offset=0.0, parent_name=None,
                 radiation_flag=10, delta_taper=0.0, **kwargs):
        kwargs.pop('_buffer', None)
        if kwargs:
            raise NameError(f'Invalid argument(s): {set(kwargs)}')
        self._parent = _parent
        self.parent_name = parent_name
        self.weight = float(weight)
        self.slice_offset = float(slice_offset)
        if radiation_flag not in (0, 10) or delta_taper != 0.0:
            raise NotImplementedError('radiation and tapering of slices are not supported')
        self.radiation_flag = int(radiation_flag)
        self.delta_taper = float(delta_taper)

    @property
    def length(self):
        return self._parent.length * self.weight if self.isthick else 0.0

    def to_dict(self):
        return {'__class__': type(self).__name__, 'weight': self.weight,
                'slice_offset': self.slice_offset, 'parent_name': self.parent_name}

    def copy(self):
        return type(self)(_parent=self._parent, weight=self.weight,
                          slice_offset=self.slice_offset, parent_name=self.parent_name)

    def raw(self):
        ptype, praw = self._parent.raw()
        return E.T_SLICE, [self.weight, float(self._kind), float(ptype)] + list(praw)

    def _transform_prefix(self):
        """The parent's frame transformations as the slice applies them
        (track_local_particle_with_transformations.h with IS_SLICE):
        length x weight for thick slices (0 for thin and edge slices),
        angle x weight, the parent's h, anchor - slice_offset; none for
        drift slices."""
        par = self._parent
        if self._kind == K_DRIFT:
            return None
        if not any(getattr(par, nn) != 0.0 for nn in E.BeamElement._ACTIVE_MISALIGN):
            return None
        curved = par._type == E.T_BEND
        if curved and self._kind in (K_THIN, K_ENTRY, K_EXIT) and (
                par.rot_x_rad != 0 or par.rot_y_rad != 0 or par.rot_s_rad_no_frame != 0):
            raise NotImplementedError(
                'rot_x_rad, rot_y_rad or rot_s_rad_no_frame on a thin slice of a curved '
                'element')
        length = par.length * self.weight if self._kind == K_THICK else 0.0
        angle = par.angle * self.weight if curved else 0.0
        h = par.h if curved else 0.0
        return [par.shift_x, par.shift_y, par.shift_s, par.rot_x_rad, par.rot_y_rad,
                par.rot_s_rad_no_frame, par.rot_shift_anchor - self.slice_offset,
                par.rot_s_rad, float(length), float(angle), float(h),
                1.0 if curved else 0.0, float(E.T_SLICE)]


def _slice_class(name, parent_class, kind):
    attrs = {'_kind': kind, '_parent_class': parent_class,
             'isthick': kind in (K_THICK, K_DRIFT),
             '_static_thick': kind in (K_THICK, K_DRIFT)}
    return type(name, (_SliceBase,), attrs)


SLICE_CLASSES = {}
for _pc in (E.Quadrupole, E.Sextupole, E.Octupole, E.Bend, E.Multipole):
    _pn = _pc.__name__
    for _prefix, _k in (('ThinSlice', K_THIN), ('ThickSlice', K_THICK),
                        ('DriftSlice', K_DRIFT)):
        SLICE_CLASSES[_prefix + _pn] = _slice_class(_prefix + _pn, _pc, _k)
    if _pc is not E.Multipole:
        SLICE_CLASSES['ThinSlice' + _pn + 'Entry'] = _slice_class(
            'ThinSlice' + _pn + 'Entry', _pc, K_ENTRY)
        SLICE_CLASSES['ThinSlice' + _pn + 'Exit'] = _slice_class(
            'ThinSlice' + _pn + 'Exit', _pc, K_EXIT)
SLICE_CLASSES['DriftSlice'] = _slice_class('DriftSlice', E.Drift, K_DRIFT)
globals().update(SLICE_CLASSES)

_EDGES = (E.Quadrupole, E.Sextupole, E.Octupole, E.Bend)


def _slice_cls(parent, prefix):
    if isinstance(parent, E.Drift):
        return SLICE_CLASSES['DriftSlice'] if prefix == 'DriftSlice' else None
    return SLICE_CLASSES.get(prefix + type(parent).__name__)


def slice_line(line, slicing_strategies, _edge_markers=True):
    """Slice `line` in place (slicing.py:215-470); returns {name: [slice
    names]}."""
    strategies = [Strategy(None, element_type=E.Drift)] + list(slicing_strategies)
    names, elems, out = [], [], {}
    registry = line.element_dict

    def add(nn, ee):
        registry[nn] = ee
        names.append(nn)
        elems.append(ee)
        return nn

    for nn, ee in zip(list(line.element_names), list(line.elements)):
        thick = getattr(ee, 'isthick', False) and getattr(ee, 'length', 0.0) != 0
        chosen = None
        if thick:
            for st in reversed(strategies):
                if st.match_element(nn, ee):
                    chosen = st.slicing
                    break
        if chosen is None or isinstance(ee, _SliceBase):
            names.append(nn)
            elems.append(ee)
            out[nn] = [nn]
            continue
        if not isinstance(ee, (E.Drift,) + _EDGES + (E.Multipole,)):
            raise NotImplementedError(f'slicing of {type(ee).__name__} is not supported')
        sub = []
        markers = _edge_markers and not isinstance(ee, E.Drift)
        if markers:
            sub.append(add(f'{nn}_entry', E.Marker()))
        if isinstance(ee, _EDGES):
            sub.append(add(f'{nn}..entry_map',
                           _slice_cls(ee, 'ThinSlice')._entry(ee, nn)))
        offset, di, ei = 0.0, 0, 0
        length = ee.length
        for w, is_drift in chosen.iter_weights(length):
            if chosen.mode == 'thin':
                if is_drift:
                    sl = _slice_cls(ee, 'DriftSlice')(_parent=ee, weight=w, slice_offset=offset,
                                                     parent_name=nn)
                    sub.append(add(f'drift_{nn}..{di}', sl))
                    di += 1
                    offset += length * w
                else:
                    tcls = _slice_cls(ee, 'ThinSlice')
                    if tcls is None:        # drifts have no thin slices
                        continue
                    sl = tcls(_parent=ee, weight=w, slice_offset=offset, parent_name=nn)
                    sub.append(add(f'{nn}..{ei}', sl))
                    ei += 1
            else:
                if isinstance(ee, E.Drift):
                    sl = SLICE_CLASSES['DriftSlice'](_parent=ee, weight=w, slice_offset=offset,
                                                    parent_name=nn)
                else:
                    sl = _slice_cls(ee, 'ThickSlice')(_parent=ee, weight=w,
                                                     slice_offset=offset, parent_name=nn)
                sub.append(add(f'{nn}..{ei}', sl))
                ei += 1
                offset += length * w
        if isinstance(ee, _EDGES):
            sub.append(add(f'{nn}..exit_map', _slice_cls(ee, 'ThinSlice')._exit(ee, nn)))
        if markers:
            sub.append(add(f'{nn}_exit', E.Marker()))
        registry.setdefault(nn, ee)
        out[nn] = sub
    line.element_names = names
    line.elements = elems
    line.discard_tracker()
    return out


def _entry(cls, parent, name):
    return SLICE_CLASSES[f'ThinSlice{type(parent).__name__}Entry'](
        _parent=parent, weight=0.0, slice_offset=0.0, parent_name=name)


def _exit(cls, parent, name):
    return SLICE_CLASSES[f'ThinSlice{type(parent).__name__}Exit'](
        _parent=parent, weight=0.0, slice_offset=parent.length, parent_name=name)


_SliceBase._entry = classmethod(_entry)
_SliceBase._exit = classmethod(_exit)
