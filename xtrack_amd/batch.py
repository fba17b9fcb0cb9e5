"""Many lattice variants tracked in one launch.

The reference tracks one line per ``Line.track`` call, one ``track_line``
launch per call (xtrack/tracker.py:1372); a study over lattice variants
(error seeds, strength scans) loops over lines.  ``BatchTracker`` puts the
particles of all variants into one particle set, each variant's particles
in a contiguous range, and tracks them in one launch: every block (64, 128
or 256 threads) holds particles of one variant only, so element dispatch
stays uniform within a wave, and blocks of longer lines are dispatched
first.  Each variant's particles end exactly as
``line.track(particles, num_turns=n)`` leaves them (full turns from
element 0, s reset at the end of each turn, no monitor).

Variants whose quadrupoles and sextupoles carry individual errors have as
many distinct element records as elements, too many for per-line literal
code.  With ``compile=True`` the batch is compiled once for all its lines
with one function per distinct element *structure* (op sequence, with the
constants read from each line's program), so that every variant runs the
same compiled code; particles with chi != 1 follow on the interpreter.

Turns run in launches of ``turns_per_launch``; with ``use_graph`` one such
launch is captured in a HIP graph (torch.cuda.CUDAGraph) and replayed.  The
autotuner (``tune``) times one turn of each (block size, waves per SIMD)
candidate on a copy of the particles and keeps the fastest: block sizes 64,
128 and 256 (smaller blocks leave fewer idle lanes in the last block of
each variant and spread a small particle set over more compute units),
occupancy 3, 4 or 5 waves per SIMD for the interpreter kernel.  The
compiled batch code is built for 5 waves per SIMD, so for a compiled batch
only the block size is searched (``waves_per_eu`` sets the compiled code's
occupancy target, 0 for the library default).  The per-thread LDS slots of the particle
state are sized by the launch (block size x 48 bytes), so every block size
runs the same code.
"""
import ctypes

import numpy as np

from . import _native
from .elements import build_raw_table
from .particles import Particles
from .tracker import TrackFlags, _LineHandle

BLOCK = 256
BLOCK_SIZES = (64, 128, 256)
# interpreter kernel: block size x occupancy; compiled code: block size
CANDIDATES = tuple((bs, w) for bs in BLOCK_SIZES for w in (3, 4, 5))
COMPILED_CANDIDATES = tuple((bs, 5) for bs in BLOCK_SIZES)


class BatchTracker:
    def __init__(self, lines, device='cuda:0', compile='auto', waves_per_eu=0):
        import torch
        self.lib = _native.load()
        if self.lib.xtm_device_count() < 1:
            raise _native.NativeError('no HIP device visible to libxtrack_amd.so')
        self.device = torch.device(device)
        if self.device.index is None:
            self.device = torch.device('cuda:0')
        self.lines = list(lines)
        if not self.lines:
            raise ValueError('BatchTracker needs at least one line')
        self.track_flags = TrackFlags()
        self._handles = []
        for ln in self.lines:
            types, offsets, params = build_raw_table(ln.elements)
            h = ctypes.c_void_p()
            _native.check(self.lib.xtm_line_create(
                types.ctypes.data, offsets.ctypes.data, params.ctypes.data,
                ctypes.c_int64(len(types)), ctypes.c_double(ln.get_length()),
                ctypes.c_int32(self.device.index), ctypes.byref(h)))
            self._handles.append(_LineHandle(self.lib, h))
        arr = (ctypes.c_void_p * len(self._handles))(*[h.ptr.value for h in self._handles])
        self._batch = ctypes.c_void_p()
        _native.check(self.lib.xtm_batch_create(arr, len(self._handles),
                                                ctypes.byref(self._batch)))
        if compile not in ('auto', True, False):
            raise ValueError("compile must be 'auto', True or False")
        if compile is not False:
            st = self.lib.xtm_batch_specialize(self._batch, int(waves_per_eu))
            # too many distinct element structures: 'auto' keeps the interpreter
            if st != 0 and not (st == _native.XTM_ERR_UNSUPPORTED and compile == 'auto'):
                _native.check(st)
        self.config = None
        self.tune_timings = None
        self.last_track_info = None

    @property
    def is_compiled(self):
        return bool(self.lib.xtm_batch_is_specialized(self._batch))

    def __del__(self):
        try:
            if getattr(self, '_batch', None) is not None and self._batch.value:
                self.lib.xtm_batch_destroy(self._batch)
                self._batch = None
        except Exception:
            pass

    # ------------------------------------------------------------------
    def _stream(self):
        import torch
        return ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def _concat(self, particles_list):
        """One device particle set holding every variant's particles, and
        the start index of each variant's range."""
        if len(particles_list) != len(self.lines):
            raise ValueError('one particle set per line is needed')
        ref = particles_list[0]
        for pp in particles_list:
            if (pp.q0, pp.mass0, pp.t_sim) != (ref.q0, ref.mass0, ref.t_sim):
                raise ValueError('all particle sets of a batch need the same q0, mass0 '
                                 'and t_sim')
        starts = np.zeros(len(particles_list) + 1, dtype=np.int64)
        starts[1:] = np.cumsum([pp._capacity for pp in particles_list])
        hosts = [pp.to_numpy() for pp in particles_list]
        big = Particles.__new__(Particles)
        big._capacity = int(starts[-1])
        big.q0, big.mass0, big.t_sim = ref.q0, ref.mass0, ref.t_sim
        big.start_tracking_at_element = -1
        big._arrays = {nn: np.concatenate([h[nn] for h in hosts]) for nn in hosts[0]}
        big._device = None
        big.move(_context=self.device)
        return big, starts

    def _split(self, big, starts, particles_list):
        host = big.to_numpy()
        for k, pp in enumerate(particles_list):
            part = {nn: np.ascontiguousarray(vv[starts[k]:starts[k + 1]])
                    for nn, vv in host.items()}
            pp._set_all(part)

    def _set_layout(self, starts, config):
        _native.check(self.lib.xtm_batch_set_layout(self._batch, starts.ctypes.data,
                                                    int(config[0]), int(config[1])))

    def _launch(self, big, num_turns):
        pst = _native.make_particles_struct(big)
        _native.check(self.lib.xtm_batch_track(
            self._batch, ctypes.byref(pst), int(num_turns),
            ctypes.c_uint64(int(self.track_flags.get_flags_register())), self._stream()))

    def blocks(self):
        """Current block table as an (n, 3) array of (line, first, count)."""
        n = self.lib.xtm_batch_num_blocks(self._batch)
        out = np.zeros((max(n, 0), 3), dtype=np.int64)
        if n > 0:
            _native.check(self.lib.xtm_batch_copy_blocks(self._batch, out.ctypes.data))
        return out

    # ------------------------------------------------------------------
    def candidates(self):
        """The (block size, waves per SIMD) configurations the tuner tries."""
        return COMPILED_CANDIDATES if self.is_compiled else CANDIDATES

    def tune(self, particles_list, turns=1, candidates=None):
        """Time `turns` turns of each candidate (block size, waves per SIMD)
        on a copy of the particles; keep the fastest as self.config."""
        import torch
        candidates = self.candidates() if candidates is None else candidates
        big, starts = self._concat(particles_list)
        timings = {}
        for cand in candidates:
            self._set_layout(starts, cand)
            work = big.copy()
            self._launch(work, 1)     # first launch: code loading
            torch.cuda.synchronize(self.device)
            work = big.copy()
            ev0 = torch.cuda.Event(enable_timing=True)
            ev1 = torch.cuda.Event(enable_timing=True)
            ev0.record()
            self._launch(work, turns)
            ev1.record()
            torch.cuda.synchronize(self.device)
            timings[cand] = ev0.elapsed_time(ev1)
        self.config = min(timings, key=timings.get)
        self.tune_timings = timings
        return self.config

    def track(self, particles_list, num_turns, turns_per_launch=None, tune=False,
              use_graph=False):
        """Track each particle set through its line for num_turns full turns,
        in place."""
        import torch
        if num_turns < 0:
            raise ValueError('num_turns must be >= 0')
        if tune and self.config is None:
            self.tune(particles_list)
        if self.config is None:
            self.config = (BLOCK, 4)
        big, starts = self._concat(particles_list)
        self._set_layout(starts, self.config)
        tpl = int(turns_per_launch or num_turns or 1)
        n_full, rest = divmod(int(num_turns), tpl)
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        graph = None
        if use_graph and n_full > 0:
            # capture one launch of tpl turns; replay it n_full times
            work = big.copy()
            self._launch(work, 1)          # code loading before capture
            torch.cuda.synchronize(self.device)
            graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(graph):
                self._launch(big, tpl)
            torch.cuda.synchronize(self.device)
        ev0.record()
        for _ in range(n_full):
            if graph is not None:
                graph.replay()
            else:
                self._launch(big, tpl)
        if rest:
            self._launch(big, rest)
        ev1.record()
        torch.cuda.synchronize(self.device)
        self.last_track_info = {'kernel_ms': ev0.elapsed_time(ev1), 'config': self.config,
                                'compiled': self.is_compiled,
                                'launches': n_full + (1 if rest else 0),
                                'graph': graph is not None,
                                'particles': int(starts[-1])}
        self._split(big, starts, particles_list)
        return self.last_track_info
