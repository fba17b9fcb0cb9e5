"""Batched lattice variants: the block table of xtm_batch_set_layout on
host-only handles (every particle in exactly one block, one line per block,
longer lines first) and the argument checks.  The tracking itself is a GPU
test (test_gpu_batch.py)."""
import ctypes

import numpy as np
import pytest

from xtrack_amd import _native, lattices
from xtrack_amd.elements import build_raw_table


def _host_batch(lines):
    lib = _native.load()
    handles = []
    for ln in lines:
        t, o, p = build_raw_table(ln.elements)
        h = ctypes.c_void_p()
        _native.check(lib.xtm_line_create(t.ctypes.data, o.ctypes.data, p.ctypes.data,
                                          len(t), ln.get_length(), -1, ctypes.byref(h)))
        handles.append(h)
    arr = (ctypes.c_void_p * len(handles))(*[h.value for h in handles])
    b = ctypes.c_void_p()
    _native.check(lib.xtm_batch_create(arr, len(handles), ctypes.byref(b)))
    return lib, handles, b


def _free(lib, handles, b):
    lib.xtm_batch_destroy(b)
    for h in handles:
        lib.xtm_line_destroy(h)


def test_block_table():
    lines = [lattices.collider_variant(s, n_arcs=1, cells_per_arc=(1, 4)) for s in range(5)]
    lib, handles, b = _host_batch(lines)
    try:
        counts = np.array([300, 0, 5, 1024, 257])
        starts = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
        _native.check(lib.xtm_batch_set_layout(b, starts.ctypes.data, 256, 4))
        n = lib.xtm_batch_num_blocks(b)
        tab = np.zeros((n, 3), dtype=np.int64)
        _native.check(lib.xtm_batch_copy_blocks(b, tab.ctypes.data))
        line, first, cnt = tab.T
        assert n == sum(-(-c // 256) for c in counts)
        assert np.all((cnt >= 1) & (cnt <= 256))
        covered = np.zeros(starts[-1], dtype=int)
        for ll, f, c in tab:
            assert starts[ll] <= f and f + c <= starts[ll + 1]
            covered[f:f + c] += 1
        assert np.all(covered == 1)
        nele = np.array([len(ln.elements) for ln in lines])
        assert np.all(np.diff(nele[line]) <= 0)          # longer lines first
        # bad layouts and kernel choices are rejected
        for bs, w in ((128, 4), (256, 2), (256, 6)):
            assert lib.xtm_batch_set_layout(b, starts.ctypes.data, bs, w) == _native.XTM_ERR_ARG
        bad = starts.copy()
        bad[2] = bad[1] - 1
        assert lib.xtm_batch_set_layout(b, bad.ctypes.data, 256, 4) == _native.XTM_ERR_ARG
        bad = starts + 1
        assert lib.xtm_batch_set_layout(b, bad.ctypes.data, 256, 4) == _native.XTM_ERR_ARG
        # a host-only batch does not track
        pst = _native.XtmParticles()
        pst.capacity = int(starts[-1])
        assert lib.xtm_batch_track(b, ctypes.byref(pst), 1, 0, None) == _native.XTM_ERR_ARG
    finally:
        _free(lib, handles, b)


def test_create_checks():
    lib = _native.load()
    b = ctypes.c_void_p()
    assert lib.xtm_batch_create(None, 0, ctypes.byref(b)) == _native.XTM_ERR_ARG
    arr = (ctypes.c_void_p * 1)(None)
    assert lib.xtm_batch_create(arr, 1, ctypes.byref(b)) == _native.XTM_ERR_ARG


def test_variant_generator():
    """cells per arc in the stated range, strengths scattered by the stated
    relative error, the same seed giving the same lattice."""
    a = lattices.collider_variant(7, n_arcs=2, cells_per_arc=(3, 6), rel_err=1e-3)
    b = lattices.collider_variant(7, n_arcs=2, cells_per_arc=(3, 6), rel_err=1e-3)
    n_cells = sum(1 for nn in a.element_names if nn.startswith('qf.a0c') and
                  not nn.endswith('.ap'))
    assert 3 <= n_cells <= 6
    k1a = np.array([ee.k1 for ee in a.elements if type(ee).__name__ == 'Quadrupole'])
    k1b = np.array([ee.k1 for ee in b.elements if type(ee).__name__ == 'Quadrupole'])
    assert np.array_equal(k1a, k1b)
    rel = np.abs(k1a) / 0.0085 - 1
    assert 1e-4 < np.std(rel) < 1e-2 and np.all(np.abs(rel) < 1e-2)
