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
    n = len(x) if fn != 7 else len(x) // 2
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
