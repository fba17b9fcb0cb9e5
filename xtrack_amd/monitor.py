"""Turn-by-turn particle monitor with the record layout of the reference
``ParticlesMonitor`` (xtrack/monitors/particles_monitor.py:12-104,
particles_monitor.h:13-77): one record per (particle id, turn) holding every
per-particle field, read back as ``monitor.<field>`` arrays of shape
(n_particles, n_turns)."""
import numpy as np

from .particles import Particles, per_particle_vars


def _zero_particles(n_records, device=None):
    out = Particles.__new__(Particles)
    out._capacity = int(n_records)
    out._arrays = {nn: np.zeros(n_records, dtype=tt) for tt, nn in per_particle_vars}
    out._device = None
    out.q0, out.mass0, out.t_sim = 1.0, 0.0, 0.0
    out.start_at_turn = 0
    out.start_tracking_at_element = -1
    if device is not None:
        out.move(_context=device)
    return out


class ParticlesMonitor:
    def __init__(self, start_at_turn=0, stop_at_turn=None, n_repetitions=None,
                 repetition_period=None, num_particles=None, particle_id_range=None,
                 auto_to_numpy=True, _context=None):
        if particle_id_range is not None:
            assert num_particles is None
            part_id_start, part_id_end = particle_id_range
        else:
            assert num_particles is not None
            part_id_start, part_id_end = 0, num_particles
        n_part_ids = int(part_id_end) - int(part_id_start)
        assert n_part_ids >= 0
        n_turns = int(stop_at_turn) - int(start_at_turn)
        if repetition_period is None:
            assert n_repetitions is None
            repetition_period = -1
            n_repetitions = 1
        self.start_at_turn = int(start_at_turn)
        self.stop_at_turn = int(stop_at_turn)
        self.part_id_start = int(part_id_start)
        self.part_id_end = int(part_id_end)
        self.ebe_mode = 0
        self.n_repetitions = int(n_repetitions)
        self.repetition_period = int(repetition_period)
        self.n_records = n_turns * n_part_ids * self.n_repetitions
        self.auto_to_numpy = auto_to_numpy
        self.data = _zero_particles(max(self.n_records, 1), device=_context)

    def _field(self, name):
        vv = self.data._host_arrays()[name] if self.auto_to_numpy else self.data._arrays[name]
        vv = vv[:self.n_records]
        n_cols = self.stop_at_turn - self.start_at_turn
        if self.n_repetitions == 1:
            return vv.reshape(self.n_records // n_cols, n_cols)
        n_rows = self.n_records // n_cols // self.n_repetitions
        return vv.reshape(self.n_repetitions, n_rows, n_cols)

    def __getattr__(self, name):
        if name in ('data', '__setstate__') or name.startswith('__'):
            raise AttributeError(name)
        if name in self.data._arrays:
            return self._field(name)
        if name == 'pzeta':
            return self._field('ptau') / self._field('beta0')
        raise AttributeError(name)
