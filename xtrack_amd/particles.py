"""SoA particle container with the field set and construction semantics of
xtrack's ``Particles`` (xtrack/particles/particles.py:26-83, constructor
:453-731, energy/reference conversions :1900-2097).

Storage is either host numpy arrays or torch tensors in HIP device memory
(``move(_context='cuda')``); the tracker consumes the device form by raw
pointer.  Construction arithmetic runs in numpy float64 with the
reference's formulas so that the derived fields (``delta/ptau/rpp/rvv``,
``beta0/gamma0``) carry the same bits.
"""
import numpy as np

from .constants import PROTON_MASS_EV, LAST_INVALID_STATE

_F8 = np.float64
_I8 = np.int64
_U4 = np.uint32

# (dtype, name) in the reference's order (particles.py:46-80)
per_particle_vars = (
    (_F8, 'p0c'), (_F8, 'gamma0'), (_F8, 'beta0'), (_F8, 's'), (_F8, 'zeta'),
    (_F8, 'x'), (_F8, 'y'), (_F8, 'px'), (_F8, 'py'),
    (_F8, 'ptau'), (_F8, 'delta'), (_F8, 'rpp'), (_F8, 'rvv'),
    (_F8, 'chi'), (_F8, 'charge_ratio'), (_F8, 'weight'), (_F8, 'ax'), (_F8, 'ay'),
    (_F8, 'spin_x'), (_F8, 'spin_y'), (_F8, 'spin_z'),
    (_F8, 'anomalous_magnetic_moment'),
    (_I8, 'pdg_id'), (_I8, 'particle_id'), (_I8, 'at_element'), (_I8, 'at_turn'),
    (_I8, 'state'), (_I8, 'parent_particle_id'),
    (_U4, '_rng_s1'), (_U4, '_rng_s2'), (_U4, '_rng_s3'), (_U4, '_rng_s4'),
)
FIELD_NAMES = tuple(nn for _, nn in per_particle_vars)
_DTYPES = {nn: tt for tt, nn in per_particle_vars}
scalar_vars = ('q0', 'mass0', 't_sim')

_accepted = set(FIELD_NAMES) | set(scalar_vars) | {
    'energy0', 'tau', 'pzeta', 'mass_ratio', 'kinetic_energy0', 'rigidity0',
    '_context', 'start_tracking_at_element', '_capacity'}


def _torch():
    import torch
    return torch


def _is_device_context(ctx):
    if ctx is None:
        return False
    if isinstance(ctx, str):
        if ctx in ('cpu', 'numpy', 'host'):
            return False
        if ctx.startswith('cuda') or ctx.startswith('hip'):
            return True
        raise ValueError(f'Unknown context {ctx!r}')
    torch = _torch()
    if isinstance(ctx, torch.device):
        return ctx.type == 'cuda'
    if hasattr(ctx, 'device'):
        return True
    raise ValueError(f'Unknown context {ctx!r}')


def _device_of(ctx):
    torch = _torch()
    if isinstance(ctx, str):
        ctx = ctx.replace('hip', 'cuda')
        return torch.device(ctx if ':' in ctx else 'cuda:0')
    if isinstance(ctx, torch.device):
        return ctx if ctx.index is not None else torch.device('cuda:0')
    return ctx.device


class Particles:
    """Particle coordinates and bookkeeping, one array slot per particle.

    Parameters follow the reference constructor: per-particle values may be
    scalars or arrays of a common length; ``p0c`` (or ``energy0``,
    ``gamma0``, ``beta0``, ``kinetic_energy0``, ``rigidity0``) fixes the
    reference, ``delta`` (or ``ptau``/``pzeta``) the energy deviation and
    ``zeta`` (or ``tau``) the longitudinal position.
    """

    per_particle_vars = per_particle_vars

    def __init__(self, _capacity=None, _context=None, **kwargs):
        unknown = set(kwargs) - _accepted
        if unknown:
            raise NameError(f'Invalid argument(s) provided: {unknown}')

        input_length = 1
        for nn in list(FIELD_NAMES) + ['energy0', 'kinetic_energy0', 'rigidity0',
                                       'tau', 'pzeta', 'mass_ratio']:
            if nn not in kwargs or kwargs[nn] is None:
                continue
            vv = kwargs[nn]
            if np.isscalar(vv) or len(vv) == 1:
                continue
            if len(vv) != input_length and input_length > 1:
                raise ValueError('All per particle vars have to be of the same length.')
            input_length = len(vv)
        if _capacity is None:
            _capacity = input_length
        if _capacity < input_length:
            raise ValueError(f'Capacity ({_capacity}) has to be greater or equal '
                             f'to the number of particles ({input_length}).')
        n = input_length
        self._capacity = int(_capacity)
        self._arrays = {}
        for tt, nn in per_particle_vars:
            fill = 0 if nn.startswith('_rng') else LAST_INVALID_STATE
            self._arrays[nn] = np.full(self._capacity, fill, dtype=tt)
        self._device = None

        def vec(name, default=None, dtype=_F8):
            vv = kwargs.get(name, default)
            if vv is None:
                return None
            vv = np.asarray(vv, dtype=dtype)
            if vv.ndim == 0 or vv.size == 1:
                return np.full(n, vv.reshape(-1)[0] if vv.ndim else vv.item(), dtype=dtype)
            return vv.copy()

        a = self._arrays
        sl = slice(0, n)
        state = vec('state', 1, _I8)
        a['state'][sl] = state
        pid = vec('particle_id', None, _I8)
        a['particle_id'][sl] = np.arange(n, dtype=_I8) if pid is None else pid
        ppid = vec('parent_particle_id', None, _I8)
        a['parent_particle_id'][sl] = a['particle_id'][sl] if ppid is None else ppid

        self.q0 = float(kwargs.get('q0', 1.0))
        self.mass0 = float(kwargs.get('mass0', PROTON_MASS_EV))
        self.t_sim = float(kwargs.get('t_sim', 0.0))
        self.start_tracking_at_element = int(kwargs.get('start_tracking_at_element', -1))

        # independent vars (particles.py:1805-1819)
        for nn, dflt in (('s', 0), ('at_turn', 0), ('at_element', 0), ('weight', 1),
                         ('x', 0), ('y', 0), ('px', 0), ('py', 0), ('ax', 0), ('ay', 0),
                         ('anomalous_magnetic_moment', 0), ('spin_x', 0), ('spin_y', 0),
                         ('spin_z', 0), ('pdg_id', 0)):
            a[nn][sl] = vec(nn, dflt, _DTYPES[nn])
        for nn in ('_rng_s1', '_rng_s2', '_rng_s3', '_rng_s4'):
            if nn in kwargs:
                a[nn][sl] = vec(nn, 0, _U4)

        mask = a['state'][sl] > LAST_INVALID_STATE
        self._init_chi_charge_ratio(kwargs, n, vec, mask)
        self._init_refs(kwargs, n, vec)
        self._init_energy_deviations(kwargs, n, vec)
        self._init_zeta(kwargs, n, vec)

    # ---- construction helpers (particles.py:1900-2097) ---------------------
    def _init_chi_charge_ratio(self, kwargs, n, vec, mask):
        a = self._arrays
        chi = vec('chi'); cr = vec('charge_ratio'); mr = vec('mass_ratio')
        nargs = sum(v is not None for v in (chi, cr, mr))
        if nargs == 0:
            a['chi'][:n] = 1.0
            a['charge_ratio'][:n] = 1.0
            return
        if nargs == 1:
            raise ValueError('Two of `chi`, `charge_ratio` and `mass_ratio` must be provided.')
        if chi is None:
            chi = cr / mr
        elif cr is None:
            cr = chi * mr
        a['chi'][:n] = chi
        a['charge_ratio'][:n] = cr

    def _init_refs(self, kwargs, n, vec):
        a = self._arrays
        m0 = self.mass0
        p0c = vec('p0c'); energy0 = vec('energy0'); gamma0 = vec('gamma0')
        beta0 = vec('beta0'); ke0 = vec('kinetic_energy0'); rig0 = vec('rigidity0')
        if all(v is None for v in (p0c, energy0, gamma0, beta0, ke0, rig0)):
            p0c = np.full(n, 1e9)
        if p0c is not None:
            _energy0 = np.sqrt(p0c ** 2 + m0 ** 2)
            _beta0 = p0c / _energy0
            _gamma0 = _energy0 / m0
            _p0c = p0c
        elif energy0 is not None:
            _p0c = np.sqrt(energy0 ** 2 - m0 ** 2)
            _beta0 = _p0c / energy0
            _gamma0 = energy0 / m0
        elif gamma0 is not None:
            _beta0 = np.sqrt(1 - 1 / gamma0 ** 2)
            _energy0 = m0 * gamma0
            _p0c = _energy0 * _beta0
            _gamma0 = gamma0
        elif beta0 is not None:
            _gamma0 = 1 / np.sqrt(1 - beta0 ** 2)
            _energy0 = m0 * _gamma0
            _p0c = _energy0 * beta0
            _beta0 = beta0
        elif ke0 is not None:
            _energy0 = ke0 + m0
            _p0c = np.sqrt(_energy0 ** 2 - m0 ** 2)
            _beta0 = _p0c / _energy0
            _gamma0 = _energy0 / m0
        else:
            from .constants import C_LIGHT
            _p0c = rig0 * abs(self.q0) * C_LIGHT
            _energy0 = np.sqrt(_p0c ** 2 + m0 ** 2)
            _beta0 = _p0c / _energy0
            _gamma0 = _energy0 / m0
        a['p0c'][:n] = p0c if p0c is not None else _p0c
        a['gamma0'][:n] = gamma0 if gamma0 is not None else _gamma0
        a['beta0'][:n] = beta0 if beta0 is not None else _beta0

    def _init_energy_deviations(self, kwargs, n, vec):
        a = self._arrays
        delta = vec('delta'); ptau = vec('ptau'); pzeta = vec('pzeta')
        if delta is None and ptau is None and pzeta is None:
            delta = np.zeros(n)
        beta0 = a['beta0'][:n]
        if delta is not None:
            _delta = delta
            _ptau = np.sqrt(_delta ** 2 + 2 * _delta + 1 / beta0 ** 2) - 1 / beta0
        elif ptau is not None:
            _ptau = ptau
            _delta = np.sqrt(_ptau ** 2 + 2 * _ptau / beta0 + 1) - 1
        else:
            _ptau = pzeta * beta0
            _delta = np.sqrt(_ptau ** 2 + 2 * _ptau / beta0 + 1) - 1
        a['delta'][:n] = _delta
        a['ptau'][:n] = _ptau
        d = a['delta'][:n]
        delta_beta0 = d * beta0
        ptau_beta0 = np.sqrt(delta_beta0 ** 2 + 2 * delta_beta0 * beta0 + 1) - 1
        rvv = vec('rvv'); rpp = vec('rpp')
        a['rvv'][:n] = rvv if rvv is not None else (1 + d) / (1 + ptau_beta0)
        a['rpp'][:n] = rpp if rpp is not None else 1 / (1 + d)

    def _init_zeta(self, kwargs, n, vec):
        a = self._arrays
        zeta = vec('zeta'); tau = vec('tau')
        if zeta is None and tau is None:
            zeta = np.zeros(n)
        a['zeta'][:n] = zeta if zeta is not None else a['beta0'][:n] * tau

    # ---- storage ------------------------------------------------------------
    @property
    def on_device(self):
        return self._device is not None

    @property
    def device(self):
        return self._device

    def move(self, _context=None):
        """Move the arrays to host numpy (``'cpu'``) or HIP device memory
        (``'cuda'``, ``'cuda:N'`` or a torch device)."""
        if _is_device_context(_context):
            torch = _torch()
            dev = _device_of(_context)
            if self._device == dev:
                return self
            host = self._host_arrays()
            self._arrays = {nn: torch.from_numpy(np.ascontiguousarray(vv)).to(dev)
                            for nn, vv in host.items()}
            self._device = dev
        else:
            if self._device is None:
                return self
            self._arrays = self._host_arrays()
            self._device = None
        return self

    def _host_arrays(self):
        if self._device is None:
            return self._arrays
        torch = _torch()
        torch.cuda.synchronize(self._device)
        return {nn: vv.cpu().numpy() for nn, vv in self._arrays.items()}

    def copy(self, _context=None):
        out = Particles.__new__(Particles)
        out._capacity = self._capacity
        out.q0, out.mass0, out.t_sim = self.q0, self.mass0, self.t_sim
        out.start_tracking_at_element = self.start_tracking_at_element
        if self._device is None:
            out._arrays = {nn: vv.copy() for nn, vv in self._arrays.items()}
        else:
            out._arrays = {nn: vv.clone() for nn, vv in self._arrays.items()}
        out._device = self._device
        if _context is not None:
            out.move(_context=_context)
        return out

    def to_numpy(self):
        """Dictionary of host numpy copies of all per-particle arrays."""
        return {nn: np.array(vv) for nn, vv in self._host_arrays().items()}

    def to_dict(self):
        out = {nn: vv.copy() for nn, vv in self.to_numpy().items()}
        out.update(q0=self.q0, mass0=self.mass0, t_sim=self.t_sim)
        return out

    @classmethod
    def from_dict(cls, dct, _context=None):
        """Rebuild from ``to_dict`` output, keeping every stored field."""
        out = cls.__new__(cls)
        arrays = {nn: np.asarray(dct[nn], dtype=tt).copy() for tt, nn in per_particle_vars}
        out._capacity = len(arrays['x'])
        out._arrays = arrays
        out._device = None
        out.q0 = float(dct.get('q0', 1.0))
        out.mass0 = float(dct.get('mass0', PROTON_MASS_EV))
        out.t_sim = float(dct.get('t_sim', 0.0))
        out.start_tracking_at_element = -1
        if _context is not None:
            out.move(_context=_context)
        return out

    # ---- reference-style helpers -------------------------------------------
    def get_active_particle_id_range(self):
        """(min, max+1) of particle ids with state > 0 (particles.py:1340-1349)."""
        h = self._host_arrays()
        ids = h['particle_id'][h['state'] > 0]
        return int(np.min(ids)), int(np.max(ids)) + 1

    def reorganize(self):
        """Active particles first, then lost, then unused slots
        (particles.py:1198-1259).  Returns (n_active, n_lost)."""
        h = {nn: np.array(vv) for nn, vv in self._host_arrays().items()}
        st = h['state']
        m_act = st > 0
        m_lost = (st < 1) & (st > LAST_INVALID_STATE)
        na, nl = int(m_act.sum()), int(m_lost.sum())
        if not m_act[:na].all():
            for tt, nn in per_particle_vars:
                if nn.startswith('_rng'):
                    continue
                vv = h[nn]
                act, lost = vv[m_act], vv[m_lost]
                vv[:na] = act
                vv[na:na + nl] = lost
                vv[na + nl:] = tt(LAST_INVALID_STATE)
            self._set_all(h)
        return na, nl

    def _set_all(self, host):
        if self._device is None:
            self._arrays = host
        else:
            torch = _torch()
            self._arrays = {nn: torch.from_numpy(vv).to(self._device) for nn, vv in host.items()}

    def sorted_by_id(self):
        """Host copy of all arrays ordered by particle_id (valid slots only)."""
        h = self.to_numpy()
        valid = h['state'] > LAST_INVALID_STATE
        order = np.argsort(h['particle_id'][valid], kind='stable')
        return {nn: vv[valid][order] for nn, vv in h.items()}

    @property
    def energy0(self):
        return np.sqrt(self._np('p0c') ** 2 + self.mass0 ** 2)

    @property
    def energy(self):
        return self.energy0 + self._np('ptau') * self._np('p0c')

    @property
    def pzeta(self):
        return self._np('ptau') / self._np('beta0')

    @property
    def tau(self):
        return self._np('zeta') / self._np('beta0')

    def _np(self, nn):
        return np.asarray(self._host_arrays()[nn])

    def __len__(self):
        return self._capacity


def _make_field_property(nn):
    def getter(self):
        return self._arrays[nn]

    def setter(self, value):
        arr = self._arrays[nn]
        if self._device is None:
            arr[...] = value
        else:
            torch = _torch()
            if isinstance(value, torch.Tensor):
                arr.copy_(value.to(arr.device, arr.dtype))
            else:
                arr.copy_(torch.as_tensor(np.asarray(value, dtype=_DTYPES[nn])).to(arr.device))
    return property(getter, setter)


for _nn in FIELD_NAMES:
    setattr(Particles, _nn, _make_field_property(_nn))
