"""CPU checker backend for the parity tests: drives the oracle
(oracle/xt_oracle.c, built by oracle/Makefile) through the same host-side
Line/Tracker code as the HIP backend.  Test infrastructure only."""
import ctypes
import os
import subprocess

import numpy as np

from xtrack_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, 'oracle')
_libs = {}


def oracle_lib(math='xm'):
    """Load oracle/_build/liboracle_<math>.so (math: 'xm' = project math
    library, 'libm' = platform libm), building it if needed."""
    if math in _libs:
        return _libs[math]
    path = os.path.join(ORACLE_DIR, '_build', f'liboracle_{math}.so')
    if not os.path.exists(path):
        subprocess.run(['make', '-C', ORACLE_DIR], check=True, capture_output=True)
    lib = ctypes.CDLL(path)
    c = ctypes
    lib.xto_track_line.restype = c.c_int32
    lib.xto_track_line.argtypes = [c.c_void_p, c.c_void_p, c.c_void_p, c.c_int64, c.c_double,
                                   c.POINTER(_native.XtmParticles), c.c_int32, c.c_int32,
                                   c.c_int32, c.c_int32, c.c_int32, c.c_int32,
                                   c.POINTER(_native.XtmMonitor), c.c_uint64, c.c_int32]
    lib.xto_math_eval.restype = c.c_int32
    lib.xto_math_eval.argtypes = [c.c_int32, c.c_void_p, c.c_void_p, c.c_int64]
    _libs[math] = lib
    return lib


def oracle_math(fn, x, math='xm'):
    lib = oracle_lib(math)
    x = np.ascontiguousarray(x, dtype=np.float64)
    n = len(x) // 2 if fn in (7, 11) else len(x)
    y = np.empty(n)
    assert lib.xto_math_eval(fn, x.ctypes.data, y.ctypes.data, n) == 0
    return y


class OracleBackend:
    """Two-method backend protocol of xtrack_amd.tracker on the CPU oracle."""
    space = 'host'

    def __init__(self, math='xm', num_threads=1):
        self.lib = oracle_lib(math)
        self.num_threads = num_threads

    def prepare_line(self, types, offsets, params, line_length):
        return (np.ascontiguousarray(types), np.ascontiguousarray(offsets),
                np.ascontiguousarray(params), float(line_length))

    def track_line(self, handle, particles, num_turns, ele_start, num_ele_track,
                   flag_end_turn_actions, flag_reset_s_at_end_turn, flag_monitor,
                   monitor, track_flags):
        types, offsets, params, line_length = handle
        pst = _native.make_particles_struct(particles)
        mst = _native.make_monitor_struct(monitor) if monitor is not None else None
        st = self.lib.xto_track_line(
            types.ctypes.data, offsets.ctypes.data, params.ctypes.data, len(types),
            line_length, ctypes.byref(pst), int(num_turns), int(ele_start),
            int(num_ele_track), int(flag_end_turn_actions), int(flag_reset_s_at_end_turn),
            int(flag_monitor), ctypes.byref(mst) if mst is not None else None,
            int(track_flags), self.num_threads)
        if st != 0:
            raise RuntimeError(f'oracle status {st}')

    def synchronize(self):
        pass

    def monitor_context(self):
        return None


_host_kernel = None


def host_kernel_lib():
    """tests/hostkernel/_build/libkernel_host.so: the kernel's element code
    compiled for the host (test infrastructure)."""
    global _host_kernel
    if _host_kernel is not None:
        return _host_kernel
    d = os.path.join(ROOT, 'tests', 'hostkernel')
    subprocess.run(['make', '-C', d], check=True, capture_output=True)
    lib = ctypes.CDLL(os.path.join(d, '_build', 'libkernel_host.so'))
    c = ctypes
    lib.xth_track_line.restype = c.c_int32
    lib.xth_track_line.argtypes = [c.c_void_p, c.c_void_p, c.c_void_p, c.c_int64, c.c_double,
                                   c.POINTER(_native.XtmParticles), c.c_int32, c.c_int32,
                                   c.c_int32, c.c_int32, c.c_int32, c.c_int32,
                                   c.POINTER(_native.XtmMonitor), c.c_uint64]
    _host_kernel = lib
    return lib


class HostKernelBackend(OracleBackend):
    """Same protocol, running the kernel's own element code on the host."""

    def __init__(self):
        self.lib = host_kernel_lib()

    def track_line(self, handle, particles, num_turns, ele_start, num_ele_track,
                   flag_end_turn_actions, flag_reset_s_at_end_turn, flag_monitor,
                   monitor, track_flags):
        types, offsets, params, line_length = handle
        pst = _native.make_particles_struct(particles)
        mst = _native.make_monitor_struct(monitor) if monitor is not None else None
        st = self.lib.xth_track_line(
            types.ctypes.data, offsets.ctypes.data, params.ctypes.data, len(types),
            line_length, ctypes.byref(pst), int(num_turns), int(ele_start),
            int(num_ele_track), int(flag_end_turn_actions), int(flag_reset_s_at_end_turn),
            int(flag_monitor), ctypes.byref(mst) if mst is not None else None,
            int(track_flags))
        if st != 0:
            raise RuntimeError(f'host kernel status {st}')


# ---------------------------------------------------------------------------
# generated line code compiled for the host (test infrastructure)
_gen_cache = {}


def generated_host_line(types, offsets, params, line_length, options=0):
    """(lib, rec, prog): the library's generated line code for this element
    table (xtm_line_generate_source_ex options), compiled for the host with
    g++, plus the record index per element and the element program it
    refers to."""
    import hashlib
    import tempfile
    lib = _native.load()
    h = ctypes.c_void_p()
    _native.check(lib.xtm_line_create(types.ctypes.data, offsets.ctypes.data,
                                      params.ctypes.data, len(types), line_length, -1,
                                      ctypes.byref(h)))
    try:
        n = ctypes.c_int64()
        rec = np.zeros(max(len(types), 1), dtype=np.int32)
        _native.check(lib.xtm_line_generate_source_ex(h, 0, options, None, 0, ctypes.byref(n),
                                                      rec.ctypes.data))
        buf = ctypes.create_string_buffer(n.value + 1)
        _native.check(lib.xtm_line_generate_source_ex(h, 0, options, buf, n.value + 1,
                                                      ctypes.byref(n), None))
        prog = np.zeros(lib.xtm_line_program_size(h), dtype=np.float64)
        _native.check(lib.xtm_line_copy_program(h, prog.ctypes.data))
        entry = np.zeros(max(len(types), 1), dtype=np.int64)
        _native.check(lib.xtm_line_copy_entry(h, entry.ctypes.data))
    finally:
        lib.xtm_line_destroy(h)
    src = buf.value.decode()
    key = hashlib.sha1(src.encode()).hexdigest()
    if key not in _gen_cache:
        d = tempfile.mkdtemp(prefix='xtm_gen_')
        cpp = os.path.join(d, 'gen.cpp')
        so = os.path.join(d, 'gen.so')
        with open(cpp, 'w') as fh:
            fh.write(src)
        subprocess.run(['g++', '-O1', '-std=c++17', '-fPIC', '-shared', '-ffp-contract=off',
                        '-fno-fast-math', f'-I{ROOT}/include',
                        f'-I{ROOT}/xtrack_amd/csrc', cpp, '-o', so],
                       check=True, capture_output=True)
        glib = ctypes.CDLL(so)
        c = ctypes
        glib.xtm_line_host.restype = None
        glib.xtm_line_host.argtypes = [c.c_void_p, c.c_void_p, c.c_void_p, c.c_double,
                                       c.POINTER(_native.XtmParticles), c.c_int32, c.c_int32,
                                       c.c_int32, c.c_int32, c.c_int32, c.c_int32,
                                       c.POINTER(_native.XtmMonitor), c.c_uint64]
        _gen_cache[key] = glib
    return _gen_cache[key], rec, prog, src, entry


class GeneratedHostBackend(OracleBackend):
    """Same protocol, running the library's generated line code on the host
    (options: xtm_line_generate_source_ex bits; with the fast arithmetic the
    host build takes exact reciprocals where the device refines hardware
    estimates)."""

    def __init__(self, options=0):
        super().__init__()
        self.options = options

    def prepare_line(self, types, offsets, params, line_length):
        glib, rec, prog, _, entry = generated_host_line(types, offsets, params, line_length,
                                                        self.options)
        return glib, rec, entry, prog, line_length

    def track_line(self, handle, particles, num_turns, ele_start, num_ele_track,
                   flag_end_turn_actions, flag_reset_s_at_end_turn, flag_monitor,
                   monitor, track_flags):
        glib, rec, entry, prog, line_length = handle
        if track_flags & 1:
            # the library runs backtracking on the interpreter
            raise NotImplementedError('generated line code tracks forward only')
        pst = _native.make_particles_struct(particles)
        mst = _native.make_monitor_struct(monitor) if monitor is not None else None
        glib.xtm_line_host(rec.ctypes.data, entry.ctypes.data, prog.ctypes.data, line_length,
                           ctypes.byref(pst),
                           int(num_turns), int(ele_start), int(num_ele_track),
                           int(flag_end_turn_actions), int(flag_reset_s_at_end_turn),
                           int(flag_monitor), ctypes.byref(mst) if mst is not None else None,
                           int(track_flags))
