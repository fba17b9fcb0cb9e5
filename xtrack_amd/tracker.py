"""Host-side tracker: turns a ``Line.track`` request into at most three
launches of the device kernel, with the segmentation, end-of-turn and
monitor semantics of the reference ``Tracker._track_no_collective``
(xtrack/tracker.py:1200-1439) and ``Tracker._track`` (:298-397).

The device work goes through a backend object; the only backend in this
package is :class:`HIPBackend` (libxtrack_amd.so).  Tests inject a CPU
checker with the same two methods to compare against.
"""
import ctypes
from time import perf_counter

import numpy as np

from . import _native
from .elements import build_raw_table
from .monitor import ParticlesMonitor

FLAG_BITS = {
    'XS_FLAG_BACKTRACK': 0,
    'XS_FLAG_KILL_CAVITY_KICK': 2,
    'XS_FLAG_IGNORE_GLOBAL_APERTURE': 3,
    'XS_FLAG_IGNORE_LOCAL_APERTURE': 4,
    'XS_FLAG_SR_TAPER': 5,
    'XS_FLAG_SR_KICK_SAME_AS_FIRST': 6,
    # no counterpart in the reference's register: freeze_longitudinal there
    # recompiles the kernel with zeta and the energy variables frozen
    'XTM_FLAG_FREEZE_LONGITUDINAL': 16,
}


class TrackFlags:
    """64-bit runtime flag register (xtrack/track_flags.py:34-47)."""

    def __init__(self, **kwargs):
        object.__setattr__(self, 'flags', {k: False for k in FLAG_BITS})
        for kk, vv in kwargs.items():
            setattr(self, kk, vv)

    def __setattr__(self, name, value):
        if name not in FLAG_BITS:
            raise AttributeError(f'Unknown track flag {name}')
        self.flags[name] = bool(value)

    def __getattr__(self, name):
        flags = object.__getattribute__(self, 'flags')
        if name in flags:
            return flags[name]
        raise AttributeError(name)

    def get_flags_register(self):
        reg = 0
        for name, bit in FLAG_BITS.items():
            if self.flags[name]:
                reg |= 1 << bit
        return np.uint64(reg)


class HIPBackend:
    """Device element table + kernel launches through libxtrack_amd.so."""
    space = 'device'

    def __init__(self, device='cuda:0', compile='auto', waves_per_eu=0, arithmetic='exact'):
        import torch
        if compile not in ('auto', True, False):
            raise ValueError("compile must be 'auto', True or False")
        if arithmetic not in ('exact', 'fast'):
            raise ValueError("arithmetic must be 'exact' or 'fast'")
        if arithmetic == 'fast' and compile is False:
            raise ValueError("arithmetic='fast' needs compiled line code (compile=True/'auto')")
        self.lib = _native.load()
        if self.lib.xtm_device_count() < 1:
            raise _native.NativeError('no HIP device visible to libxtrack_amd.so')
        self.device = torch.device(device)
        if self.device.index is None:
            self.device = torch.device('cuda:0')
        self.compile = compile
        self.waves_per_eu = int(waves_per_eu)
        self.arithmetic = arithmetic

    def prepare_line(self, types, offsets, params, line_length):
        handle = ctypes.c_void_p()
        _native.check(self.lib.xtm_line_create(
            types.ctypes.data, offsets.ctypes.data, params.ctypes.data,
            ctypes.c_int64(len(types)), ctypes.c_double(line_length),
            ctypes.c_int32(self.device.index), ctypes.byref(handle)))
        lh = _LineHandle(self.lib, handle)
        if self.compile is not False:
            opts = _native.XTM_SPEC_FAST_ARITH if self.arithmetic == 'fast' else 0
            st = self.lib.xtm_line_specialize_ex(handle, self.waves_per_eu, opts)
            # too many distinct records: 'auto' keeps the interpreter
            if st != 0 and not (st == _native.XTM_ERR_UNSUPPORTED and self.compile == 'auto'):
                _native.check(st)
        return lh

    def is_compiled(self, handle):
        return bool(self.lib.xtm_line_is_specialized(handle.ptr))

    def stream(self):
        import torch
        return ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def track_line(self, handle, particles, num_turns, ele_start, num_ele_track,
                   flag_end_turn_actions, flag_reset_s_at_end_turn, flag_monitor,
                   monitor, track_flags):
        pst = _native.make_particles_struct(particles)
        mst = _native.make_monitor_struct(monitor) if monitor is not None else None
        _native.check(self.lib.xtm_track_line(
            handle.ptr, ctypes.byref(pst), int(num_turns), int(ele_start),
            int(num_ele_track), int(flag_end_turn_actions), int(flag_reset_s_at_end_turn),
            int(flag_monitor), ctypes.byref(mst) if mst is not None else None,
            ctypes.c_uint64(int(track_flags)), self.stream()))

    def synchronize(self):
        _native.check(self.lib.xtm_synchronize(self.stream()))

    def monitor_context(self):
        return self.device


class _LineHandle:
    def __init__(self, lib, ptr):
        self.lib = lib
        self.ptr = ptr

    @property
    def num_records(self):
        return self.lib.xtm_line_num_records(self.ptr)

    def __del__(self):
        try:
            if self.ptr:
                self.lib.xtm_line_destroy(self.ptr)
                self.ptr = None
        except Exception:
            pass


class Tracker:
    def __init__(self, line, backend=None, device='cuda:0', compile='auto', waves_per_eu=0,
                 arithmetic='exact'):
        self.line = line
        self.backend = (backend if backend is not None
                        else HIPBackend(device, compile=compile, waves_per_eu=waves_per_eu,
                                        arithmetic=arithmetic))
        self.track_flags = TrackFlags()
        self.num_elements = len(line.elements)
        self.line_length = line.get_length()
        types, offsets, params = build_raw_table(line.elements)
        self._raw = (types, offsets, params)
        self._handle = self.backend.prepare_line(types, offsets, params, self.line_length)
        self.record_last_track = None
        self.time_last_track = None

    # -- particle placement ---------------------------------------------------
    def _place(self, particles):
        """Return (particles as the backend needs them, restore callback)."""
        if self.backend.space == 'device':
            if particles.on_device:
                return particles, None
            dev = self.backend.monitor_context()
            particles.move(_context=dev)
            return particles, lambda: particles.move(_context='cpu')
        if particles.on_device:
            dev = particles.device
            particles.move(_context='cpu')
            return particles, lambda: particles.move(_context=dev)
        return particles, None

    def _monitor_context(self):
        if self.backend.space == 'device':
            return self.backend.monitor_context()
        return None

    def _get_monitor(self, particles, turn_by_turn_monitor, num_turns):
        """tracker.py:1445-1480"""
        if turn_by_turn_monitor is None or turn_by_turn_monitor is False:
            return 0, None
        if turn_by_turn_monitor is True:
            mon = ParticlesMonitor(start_at_turn=0, stop_at_turn=num_turns,
                                   particle_id_range=particles.get_active_particle_id_range(),
                                   _context=self._monitor_context())
            return 1, mon
        if turn_by_turn_monitor == 'ONE_TURN_EBE':
            _, mon = self._get_monitor(particles, True, self.num_elements + 1)
            mon.ebe_mode = 1
            return 2, mon
        if isinstance(turn_by_turn_monitor, ParticlesMonitor):
            mon = turn_by_turn_monitor
            if self.backend.space == 'device' and not mon.data.on_device:
                mon.data.move(_context=self._monitor_context())
            return (2 if mon.ebe_mode == 1 else 1), mon
        raise ValueError('Please provide a valid monitor object')

    # -- public entry ---------------------------------------------------------
    def _track(self, particles, ele_start=0, ele_stop=None, num_elements=None,
               num_turns=None, turn_by_turn_monitor=None, freeze_longitudinal=False,
               backtrack=False, time=False, with_progress=False):
        if backtrack is not False and backtrack is not True and backtrack != 'force':
            raise ValueError("backtrack must be True, False or 'force'")
        if freeze_longitudinal:
            self._check_freeze_supported()
        t0 = perf_counter() if time else None
        particles, restore = self._place(particles)
        # tracker.py:1222-1236: backtracking sets XS_FLAG_BACKTRACK for this
        # call (every element of this build has a backtracking map)
        old_bt = self.track_flags.XS_FLAG_BACKTRACK
        old_fz = self.track_flags.XTM_FLAG_FREEZE_LONGITUDINAL
        if backtrack is not False:
            self.track_flags.XS_FLAG_BACKTRACK = True
        # tracker.py:1237-1242: zeta and the energy variables stay as they are
        if freeze_longitudinal:
            self.track_flags.XTM_FLAG_FREEZE_LONGITUDINAL = True
        try:
            if with_progress:
                self._track_in_batches(particles, with_progress, ele_start, ele_stop,
                                       num_elements, num_turns, turn_by_turn_monitor)
            else:
                self._track_no_collective(particles, ele_start=ele_start, ele_stop=ele_stop,
                                          num_elements=num_elements, num_turns=num_turns,
                                          turn_by_turn_monitor=turn_by_turn_monitor)
            if time:
                self.backend.synchronize()
        finally:
            self.track_flags.XS_FLAG_BACKTRACK = old_bt
            self.track_flags.XTM_FLAG_FREEZE_LONGITUDINAL = old_fz
            if restore is not None:
                restore()
        self.time_last_track = (perf_counter() - t0) if time else None

    def _check_freeze_supported(self):
        """The kernel freezes zeta by restoring it after each element, which
        equals the reference's frozen setters unless an element reads zeta
        after changing it: only an RF element (Cavity, RFMultipole) whose
        misalignment rotates about x or y or shifts along s does."""
        for nn, ee in zip(self.line.element_names, self.line.elements):
            if type(ee).__name__ in ('Cavity', 'RFMultipole') and any(
                    getattr(ee, a, 0.0) != 0.0 for a in ('rot_x_rad', 'rot_y_rad', 'shift_s')):
                raise NotImplementedError(
                    f'freeze_longitudinal with {nn}: an RF element rotated about x or y or '
                    'shifted along s is not supported')

    def _track_in_batches(self, particles, with_progress, ele_start, ele_stop,
                          num_elements, num_turns, turn_by_turn_monitor):
        """Turn batching of Tracker._track (tracker.py:313-381)."""
        if num_turns is None:
            raise ValueError('Tracking with progress indicator is only possible over '
                             'more than one turn.')
        batch_size = 100 if with_progress is True else int(with_progress)
        es = ele_start or 0
        st = ele_stop if ele_stop is not None else self.num_elements
        if isinstance(es, str):
            es = self.line.element_names.index(es)
        if isinstance(st, str):
            st = self.line.element_names.index(st)
        if es >= st:
            num_turns += 1
        if turn_by_turn_monitor is True:
            _, turn_by_turn_monitor = self._get_monitor(particles, True, num_turns)
        for ii in range(0, num_turns, batch_size):
            kw = dict(ele_start=ele_start, ele_stop=ele_stop, num_elements=num_elements,
                      num_turns=num_turns, turn_by_turn_monitor=turn_by_turn_monitor)
            first = ii == 0
            last = ii + batch_size >= num_turns
            if first and last:
                pass
            elif first:
                kw['ele_stop'] = None
                kw['num_turns'] = batch_size
            elif last:
                kw['ele_start'] = None
                rem = num_turns % batch_size
                kw['num_turns'] = rem if rem else batch_size
            else:
                kw['num_turns'] = batch_size
                kw['ele_start'] = None
                kw['ele_stop'] = None
            self._track_no_collective(particles, **kw)

    def _track_no_collective(self, particles, ele_start=None, ele_stop=None,
                             num_elements=None, num_turns=None, turn_by_turn_monitor=None):
        """tracker.py:1252-1436: split into first / middle / last launches."""
        n_el = self.num_elements
        if particles.start_tracking_at_element >= 0:
            if ele_start not in (0, None):
                raise ValueError('The argument ele_start is used, but '
                                 'particles.start_tracking_at_element is set as well. '
                                 'Please use only one of those methods.')
            ele_start = particles.start_tracking_at_element
            particles.start_tracking_at_element = -1
        if isinstance(ele_start, str):
            ele_start = self.line.element_names.index(ele_start)
        if ele_start is None:
            ele_start = 0
        assert 0 <= ele_start <= n_el
        track_flags = self.track_flags.get_flags_register()

        num_middle_turns = 0
        num_elements_last_turn = 0
        if num_elements is not None:
            assert num_elements >= 0
            if ele_stop is not None:
                raise ValueError('Cannot use both num_elements and ele_stop!')
            if num_turns is not None:
                raise ValueError('Cannot use both num_elements and num_turns!')
            if num_elements + ele_start <= n_el:
                num_elements_first_turn = num_elements
            else:
                num_elements_first_turn = n_el - ele_start
                num_middle_turns, ele_stop = divmod(ele_start + num_elements, n_el)
                num_elements_last_turn = ele_stop
                num_middle_turns -= 1
        else:
            if num_turns is None:
                num_turns = 1
            else:
                assert num_turns > 0
            if ele_stop is None:
                num_elements_first_turn = n_el - ele_start
                num_middle_turns = num_turns - 1
            else:
                if isinstance(ele_stop, str):
                    ele_stop = self.line.element_names.index(ele_stop)
                assert 0 <= ele_stop <= n_el
                if ele_stop <= ele_start:
                    num_turns += 1
                if num_turns == 1:
                    num_elements_first_turn = ele_stop - ele_start
                else:
                    num_elements_first_turn = n_el - ele_start
                    num_middle_turns = num_turns - 2
                    num_elements_last_turn = ele_stop

        if self.line.skip_end_turn_actions:
            end_first = False
            end_middle = False
        else:
            end_first = (num_elements_first_turn + ele_start == n_el)
            end_middle = True

        monitor_turns = num_middle_turns + (2 if num_elements_last_turn > 0 else 1)
        flag_monitor, monitor = self._get_monitor(particles, turn_by_turn_monitor,
                                                  monitor_turns)
        reset_s = self.line.reset_s_at_end_turn
        h = self._handle
        b = self.backend
        assert num_elements_first_turn >= 0
        b.track_line(h, particles, 1, ele_start, num_elements_first_turn, end_first,
                     reset_s, flag_monitor, monitor, track_flags)
        if num_middle_turns > 0:
            assert n_el > 0
            b.track_line(h, particles, num_middle_turns, 0, n_el, end_middle, reset_s,
                         flag_monitor, monitor, track_flags)
        if num_elements_last_turn > 0:
            b.track_line(h, particles, 1, 0, num_elements_last_turn, False, reset_s,
                         flag_monitor, monitor, track_flags)
        self.record_last_track = monitor
