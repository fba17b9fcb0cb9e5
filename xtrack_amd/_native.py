"""ctypes binding of libxtrack_amd.so (declarations in include/xtrack_amd.h).

The library is loaded from the package directory, where build() places it.
There is no fallback: if the library is missing or the HIP runtime cannot be
initialised, every tracking call raises.
"""
import ctypes
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_NAME = 'libxtrack_amd.so'

_F8P = ctypes.POINTER(ctypes.c_double)
_I8P = ctypes.POINTER(ctypes.c_int64)
_U4P = ctypes.POINTER(ctypes.c_uint32)

PARTICLE_FIELDS = (
    ('p0c', _F8P), ('gamma0', _F8P), ('beta0', _F8P), ('s', _F8P), ('zeta', _F8P),
    ('x', _F8P), ('y', _F8P), ('px', _F8P), ('py', _F8P),
    ('ptau', _F8P), ('delta', _F8P), ('rpp', _F8P), ('rvv', _F8P),
    ('chi', _F8P), ('charge_ratio', _F8P), ('weight', _F8P), ('ax', _F8P), ('ay', _F8P),
    ('spin_x', _F8P), ('spin_y', _F8P), ('spin_z', _F8P),
    ('anomalous_magnetic_moment', _F8P),
    ('pdg_id', _I8P), ('particle_id', _I8P), ('at_element', _I8P), ('at_turn', _I8P),
    ('state', _I8P), ('parent_particle_id', _I8P),
    ('rng_s1', _U4P), ('rng_s2', _U4P), ('rng_s3', _U4P), ('rng_s4', _U4P),
)


class XtmParticles(ctypes.Structure):
    _fields_ = [('capacity', ctypes.c_int64), ('q0', ctypes.c_double),
                ('mass0', ctypes.c_double), ('t_sim', ctypes.c_double)] + list(PARTICLE_FIELDS)


class XtmMonitor(ctypes.Structure):
    _fields_ = [('start_at_turn', ctypes.c_int64), ('stop_at_turn', ctypes.c_int64),
                ('part_id_start', ctypes.c_int64), ('part_id_end', ctypes.c_int64),
                ('ebe_mode', ctypes.c_int64), ('n_records', ctypes.c_int64),
                ('n_repetitions', ctypes.c_int64), ('repetition_period', ctypes.c_int64),
                ('data', XtmParticles)]


XTM_FEAT_N = 17


class XtmFeatures(ctypes.Structure):
    _fields_ = [('acc', ctypes.c_void_p), ('capacity', ctypes.c_int64),
                ('start_at_turn', ctypes.c_int64), ('stop_at_turn', ctypes.c_int64),
                ('inv_w', ctypes.c_double * 16), ('co', ctypes.c_double * 4),
                ('inv_sigma', ctypes.c_double * 2)]


def make_features_struct(acc, capacity, start_at_turn, stop_at_turn, inv_w, co, inv_sigma):
    """XtmFeatures over the accumulator array `acc` (XTM_FEAT_N x capacity
    doubles, device memory for the GPU)."""
    st = XtmFeatures()
    st.acc = _ptr(acc)
    st.capacity = int(capacity)
    st.start_at_turn, st.stop_at_turn = int(start_at_turn), int(stop_at_turn)
    w = [float(v) for v in list(inv_w.reshape(-1))]
    if len(w) != 16:
        raise ValueError('inv_w must be 4 x 4')
    st.inv_w[:] = w
    st.co[:] = [float(v) for v in co]
    st.inv_sigma[:] = [float(v) for v in inv_sigma]
    return st


def _ptr(arr):
    """Raw data pointer of a numpy array or torch tensor."""
    if hasattr(arr, 'data_ptr'):
        return arr.data_ptr()
    return arr.ctypes.data


def make_particles_struct(particles):
    """XtmParticles pointing at the arrays of a Particles object (host or
    device, whichever they live in)."""
    st = XtmParticles()
    st.capacity = particles._capacity
    st.q0 = particles.q0
    st.mass0 = particles.mass0
    st.t_sim = particles.t_sim
    arrs = particles._arrays
    for nn, tp in PARTICLE_FIELDS:
        key = '_' + nn if nn.startswith('rng') else nn
        setattr(st, nn, ctypes.cast(ctypes.c_void_p(_ptr(arrs[key])), tp))
    return st


def make_monitor_struct(monitor):
    st = XtmMonitor()
    st.start_at_turn = monitor.start_at_turn
    st.stop_at_turn = monitor.stop_at_turn
    st.part_id_start = monitor.part_id_start
    st.part_id_end = monitor.part_id_end
    st.ebe_mode = monitor.ebe_mode
    st.n_records = monitor.n_records
    st.n_repetitions = monitor.n_repetitions
    st.repetition_period = monitor.repetition_period
    st.data = make_particles_struct(monitor.data)
    return st


XTM_ERR_ARG = -1
XTM_ERR_UNSUPPORTED = -2
XTM_SPEC_FAST_ARITH = 1   # xtm_line_specialize_ex option bit
XTM_ERR_HIP = -3


class NativeError(RuntimeError):
    pass


_lib = None


def lib_path():
    """Package copy of the library, or the file named by XTRACK_AMD_LIB."""
    return os.environ.get('XTRACK_AMD_LIB') or os.path.join(_PKG, LIB_NAME)


def load():
    """Load libxtrack_amd.so once; raise NativeError if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    path = lib_path()
    if not os.path.exists(path):
        raise NativeError(f'{path} not found: build the HIP library first '
                          '(python -c "import __graft_entry__ as g; g.build()")')
    lib = ctypes.CDLL(path)
    c = ctypes
    lib.xtm_abi_version.restype = c.c_int32
    lib.xtm_last_error.restype = c.c_char_p
    lib.xtm_device_count.restype = c.c_int32
    lib.xtm_line_create.restype = c.c_int32
    lib.xtm_line_create.argtypes = [c.c_void_p, c.c_void_p, c.c_void_p, c.c_int64,
                                    c.c_double, c.c_int32, c.POINTER(c.c_void_p)]
    lib.xtm_line_destroy.restype = c.c_int32
    lib.xtm_line_destroy.argtypes = [c.c_void_p]
    lib.xtm_line_num_elements.restype = c.c_int64
    lib.xtm_line_num_elements.argtypes = [c.c_void_p]
    lib.xtm_line_num_records.restype = c.c_int64
    lib.xtm_line_num_records.argtypes = [c.c_void_p]
    lib.xtm_track_line.restype = c.c_int32
    lib.xtm_track_line.argtypes = [c.c_void_p, c.POINTER(XtmParticles), c.c_int32, c.c_int32,
                                   c.c_int32, c.c_int32, c.c_int32, c.c_int32,
                                   c.POINTER(XtmMonitor), c.c_uint64, c.c_void_p]
    lib.xtm_features_record.restype = c.c_int32
    lib.xtm_features_record.argtypes = [c.POINTER(XtmParticles), c.POINTER(XtmFeatures),
                                        c.c_void_p]
    lib.xtm_track_line_features.restype = c.c_int32
    lib.xtm_track_line_features.argtypes = [c.c_void_p, c.POINTER(XtmParticles), c.c_int32,
                                            c.POINTER(XtmFeatures), c.c_uint64, c.c_void_p]
    lib.xtm_line_count_ops.restype = c.c_int32
    lib.xtm_line_count_ops.argtypes = [c.c_void_p, c.c_void_p, c.c_double, c.c_double,
                                       c.c_double, c.c_void_p]
    lib.xtm_line_specialize.restype = c.c_int32
    lib.xtm_line_specialize.argtypes = [c.c_void_p, c.c_int32]
    lib.xtm_line_is_specialized.restype = c.c_int32
    lib.xtm_line_is_specialized.argtypes = [c.c_void_p]
    lib.xtm_line_specialize_ex.restype = c.c_int32
    lib.xtm_line_specialize_ex.argtypes = [c.c_void_p, c.c_int32, c.c_uint32]
    lib.xtm_line_specialized_options.restype = c.c_int32
    lib.xtm_line_specialized_options.argtypes = [c.c_void_p]
    lib.xtm_line_generate_source_ex.restype = c.c_int32
    lib.xtm_line_generate_source_ex.argtypes = [c.c_void_p, c.c_int32, c.c_uint32, c.c_char_p,
                                                c.c_int64, c.POINTER(c.c_int64), c.c_void_p]
    lib.xtm_line_generate_source.restype = c.c_int32
    lib.xtm_line_generate_source.argtypes = [c.c_void_p, c.c_int32, c.c_char_p, c.c_int64,
                                             c.POINTER(c.c_int64), c.c_void_p]
    lib.xtm_line_program_size.restype = c.c_int64
    lib.xtm_line_program_size.argtypes = [c.c_void_p]
    lib.xtm_line_copy_program.restype = c.c_int32
    lib.xtm_line_copy_program.argtypes = [c.c_void_p, c.c_void_p]
    lib.xtm_line_copy_entry.restype = c.c_int32
    lib.xtm_line_copy_entry.argtypes = [c.c_void_p, c.c_void_p]
    lib.xtm_batch_create.restype = c.c_int32
    lib.xtm_batch_create.argtypes = [c.c_void_p, c.c_int32, c.POINTER(c.c_void_p)]
    lib.xtm_batch_destroy.restype = c.c_int32
    lib.xtm_batch_destroy.argtypes = [c.c_void_p]
    lib.xtm_batch_set_layout.restype = c.c_int32
    lib.xtm_batch_set_layout.argtypes = [c.c_void_p, c.c_void_p, c.c_int32, c.c_int32]
    lib.xtm_batch_num_blocks.restype = c.c_int64
    lib.xtm_batch_num_blocks.argtypes = [c.c_void_p]
    lib.xtm_batch_copy_blocks.restype = c.c_int32
    lib.xtm_batch_copy_blocks.argtypes = [c.c_void_p, c.c_void_p]
    lib.xtm_batch_track.restype = c.c_int32
    lib.xtm_batch_track.argtypes = [c.c_void_p, c.POINTER(XtmParticles), c.c_int32,
                                    c.c_uint64, c.c_void_p]
    lib.xtm_batch_track_features.restype = c.c_int32
    lib.xtm_batch_track_features.argtypes = [c.c_void_p, c.POINTER(XtmParticles), c.c_int32,
                                             c.POINTER(XtmFeatures), c.c_uint64, c.c_void_p]
    lib.xtm_batch_specialize.restype = c.c_int32
    lib.xtm_batch_specialize.argtypes = [c.c_void_p, c.c_int32]
    lib.xtm_batch_is_specialized.restype = c.c_int32
    lib.xtm_batch_is_specialized.argtypes = [c.c_void_p]
    lib.xtm_batch_generate_source.restype = c.c_int32
    lib.xtm_batch_generate_source.argtypes = [c.c_void_p, c.c_int32, c.c_char_p, c.c_int64,
                                              c.POINTER(c.c_int64), c.c_void_p]
    lib.xtm_gemm_bf16.restype = c.c_int32
    lib.xtm_gemm_bf16.argtypes = [c.c_void_p, c.c_int64, c.c_int64, c.c_void_p, c.c_int64,
                                  c.c_int64, c.c_void_p, c.c_int64, c.c_void_p, c.c_void_p,
                                  c.c_int32, c.c_int64, c.c_int64, c.c_int64, c.c_void_p]
    lib.xtm_gemm_bf16_noalloc.restype = c.c_int32
    lib.xtm_gemm_bf16_noalloc.argtypes = lib.xtm_gemm_bf16.argtypes
    lib.xtm_gemm_bf16_scratch_floats.restype = c.c_int64
    lib.xtm_synchronize.restype = c.c_int32
    lib.xtm_synchronize.argtypes = [c.c_void_p]
    lib.xtm_math_eval.restype = c.c_int32
    lib.xtm_math_eval.argtypes = [c.c_int32, c.c_void_p, c.c_void_p, c.c_int64, c.c_void_p]
    if lib.xtm_abi_version() != 1:
        raise NativeError('libxtrack_amd.so ABI version mismatch')
    _lib = lib
    return lib


def check(status):
    if status != 0:
        msg = load().xtm_last_error().decode()
        raise NativeError(f'libxtrack_amd status {status}: {msg}')
