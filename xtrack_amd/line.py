"""``Line``: an ordered list of named elements plus the tracking front door
``Line.track`` (xtrack/line.py:1902-1993) and the default configuration of
the reference (line.py:283-299): no synchrotron radiation, global |x|,|y|
limit of 1 m after statically thick elements, s reset at the end of each
turn."""
from .elements import ELEMENT_CLASSES
from .particles import Particles
from .tracker import Tracker


class Line:
    def __init__(self, elements=(), element_names=None, particle_ref=None):
        if isinstance(elements, dict):
            element_names = list(elements.keys())
            elements = list(elements.values())
        elements = list(elements)
        if element_names is None:
            element_names = [f'e{ii}' for ii in range(len(elements))]
        if len(element_names) != len(elements):
            raise ValueError('elements and element_names differ in length')
        self.elements = elements
        self.element_names = list(element_names)
        # name -> element, including elements referred to but not in the
        # sequence (the parents of slices)
        self.element_dict = dict(zip(self.element_names, self.elements))
        self.particle_ref = particle_ref
        self.tracker = None
        self.reset_s_at_end_turn = True
        self.skip_end_turn_actions = False

    def __len__(self):
        return len(self.elements)

    def __getitem__(self, key):
        if isinstance(key, str):
            return self.elements[self.element_names.index(key)]
        return self.elements[key]

    def get_length(self):
        """Sum of lengths of thick elements."""
        tot = 0.0
        for ee in self.elements:
            if ee.isthick:
                tot += getattr(ee, 'length', 0.0)
        return tot

    def build_tracker(self, _context=None, backend=None, device=None, compile='auto',
                      waves_per_eu=0, arithmetic='exact'):
        """Build the device element table.  ``_context``/``device`` pick the
        HIP device (default ``cuda:0``); ``backend`` substitutes another
        implementation of the two-method backend protocol.

        ``compile``: generate and compile line code for this line (the
        counterpart of the reference compiling its tracking kernel at this
        point): 'auto' compiles when the line has few enough distinct
        element records and otherwise uses the element-program interpreter,
        True requires compiled code, False always interprets.
        ``waves_per_eu``: occupancy target of the compiled kernel (0: the
        library default).
        ``arithmetic``: 'exact' (default) evaluates every quotient and square
        root as the correctly rounded IEEE operation, in the reference's
        order, so coordinates match the reference's arithmetic bit for bit
        given the same elementary functions; 'fast' (compiled code only)
        re-associates the sector-bend polar drift onto reciprocals with a few
        ulp of error per element.  Particles with chi != 1 and backtracking
        always run the exact interpreter."""
        dev = device or (_context if isinstance(_context, str) else None) or 'cuda:0'
        self.tracker = Tracker(self, backend=backend, device=dev, compile=compile,
                               waves_per_eu=waves_per_eu, arithmetic=arithmetic)
        return self.tracker

    def discard_tracker(self):
        self.tracker = None

    @property
    def record_last_track(self):
        return self.tracker.record_last_track if self.tracker else None

    @property
    def time_last_track(self):
        return self.tracker.time_last_track if self.tracker else None

    @property
    def track_flags(self):
        return self.tracker.track_flags

    def track(self, particles, ele_start=0, ele_stop=None, num_elements=None,
              num_turns=None, turn_by_turn_monitor=None, freeze_longitudinal=False,
              time=False, with_progress=False, backtrack=False):
        """Track ``particles`` in place (semantics of xtrack Line.track)."""
        if self.tracker is None:
            self.build_tracker()
        return self.tracker._track(particles, ele_start=ele_start, ele_stop=ele_stop,
                                   num_elements=num_elements, num_turns=num_turns,
                                   turn_by_turn_monitor=turn_by_turn_monitor,
                                   freeze_longitudinal=freeze_longitudinal,
                                   backtrack=backtrack, time=time,
                                   with_progress=with_progress)

    def twiss(self, particle_ref=None, method='6d', delta0=0.0, **kwargs):
        """Linear optics of the ring from its tracker (xtrack line.py
        Line.twiss); see xtrack_amd.twiss.twiss."""
        from .twiss import twiss as _twiss
        return _twiss(self, particle_ref=particle_ref, method=method, delta0=delta0, **kwargs)

    def build_particles(self, **kwargs):
        """Particles with the reference quantities of ``particle_ref`` and the
        given physical coordinates."""
        ref = self.particle_ref
        if ref is None:
            raise ValueError('line.particle_ref is not set')
        kw = dict(q0=ref.q0, mass0=ref.mass0, p0c=float(ref._np('p0c')[0]))
        kw.update(kwargs)
        return Particles(**kw)

    def to_dict(self):
        return {'elements': {nn: ee.to_dict() for nn, ee in self.element_dict.items()},
                'element_names': list(self.element_names)}

    def slice_thick_elements(self, slicing_strategies):
        """Replace thick elements by slices in place (xtrack line.py
        Line.slice_thick_elements, slicing.py:215-470); returns
        {name: [slice names]}."""
        from .slicing import slice_line
        return slice_line(self, slicing_strategies)

    @classmethod
    def from_dict(cls, dct):
        """Line from the reference's serialised form (line.py:384,
        Line.to_dict / json): `elements` (name -> element dict with
        `__class__`) and `element_names`; a `particle_ref` dict, when
        present, becomes `line.particle_ref`.  Element classes outside this
        tracker raise NotImplementedError."""
        from .elements import _METADATA_KEYS
        from .slicing import SLICE_CLASSES
        names = dct.get('element_names') or list(dct['elements'].keys())
        built = {}

        def build(nn):
            if nn in built:                    # a name may repeat in the sequence
                return built[nn]
            ed = dict(dct['elements'][nn])
            cls_name = ed.pop('__class__')
            if cls_name in SLICE_CLASSES:
                # slices refer to their parent by name (slice_base.py:34-44)
                parent_name = ed.pop('parent_name')
                for kk in _METADATA_KEYS:
                    ed.pop(kk, None)
                built[nn] = SLICE_CLASSES[cls_name](_parent=build(parent_name),
                                                    parent_name=parent_name, **ed)
                return built[nn]
            if cls_name not in ELEMENT_CLASSES:
                raise NotImplementedError(f'element class {cls_name} ({nn}) is not supported')
            for kk in _METADATA_KEYS:
                ed.pop(kk, None)
            built[nn] = ELEMENT_CLASSES[cls_name](**ed)
            return built[nn]

        elems = [build(nn) for nn in names]
        line = cls(elements=elems, element_names=names)
        line.element_dict.update(built)
        if dct.get('particle_ref'):
            line.particle_ref = Particles.from_dict(dct['particle_ref'])
        return line

    @classmethod
    def from_json(cls, path):
        """Load a line written by the reference's Line.to_json (plain or
        gzip-compressed); a file holding {'line': ..., 'particle': ...} is
        accepted too (the particle is then available as `line.json_particle`)."""
        import gzip
        import json
        opener = gzip.open if str(path).endswith('.gz') else open
        with opener(path, 'rt') as fh:
            dct = json.load(fh)
        particle = None
        if 'line' in dct and 'elements' not in dct:
            particle = dct.get('particle')
            dct = dct['line']
        line = cls.from_dict(dct)
        line.json_particle = Particles.from_dict(particle) if particle else None
        return line
