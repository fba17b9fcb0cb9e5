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
        nele = np.array([len(ln.elements) for ln in lines])
        for bs in (64, 128, 256):
            _native.check(lib.xtm_batch_set_layout(b, starts.ctypes.data, bs, 4))
            n = lib.xtm_batch_num_blocks(b)
            tab = np.zeros((n, 3), dtype=np.int64)
            _native.check(lib.xtm_batch_copy_blocks(b, tab.ctypes.data))
            line, first, cnt = tab.T
            assert n == sum(-(-c // bs) for c in counts)
            assert np.all((cnt >= 1) & (cnt <= bs))
            covered = np.zeros(starts[-1], dtype=int)
            for ll, f, c in tab:
                assert starts[ll] <= f and f + c <= starts[ll + 1]
                covered[f:f + c] += 1
            assert np.all(covered == 1)
            assert np.all(np.diff(nele[line]) <= 0)          # longer lines first
        # bad layouts and kernel choices are rejected
        for bs, w in ((96, 4), (32, 4), (512, 4), (256, 2), (256, 6)):
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


def test_structural_batch_code_host_vs_oracle():
    """The batch source (one function per element structure, constants read
    from each line's program) compiled for the host: every variant's
    particles equal the oracle's for that variant alone, and variants that
    differ only in strengths share all their structures."""
    import os
    import subprocess
    import tempfile
    import xtrack_amd as xt
    from helpers import mixed_line, random_particles, compare_exact
    from oracle_backend import OracleBackend, ROOT
    mk = [lambda: lattices.collider_variant(1, n_arcs=8, cells_per_arc=(2, 2), rel_err=1e-3),
          lambda: lattices.collider_variant(2, n_arcs=8, cells_per_arc=(2, 2), rel_err=1e-3),
          lambda: mixed_line(n_cells=2)]
    lines = [f() for f in mk]
    lib, handles, b = _host_batch(lines)
    try:
        n = ctypes.c_int64()
        nele = np.array([len(ln.elements) for ln in lines], dtype=np.int64)
        recs = np.zeros(nele.sum(), dtype=np.int32)
        _native.check(lib.xtm_batch_generate_source(b, 0, None, 0, ctypes.byref(n),
                                                    recs.ctypes.data))
        buf = ctypes.create_string_buffer(n.value + 1)
        _native.check(lib.xtm_batch_generate_source(b, 0, buf, n.value + 1, ctypes.byref(n),
                                                    None))
        src = buf.value.decode()
        entries, progs = [], []
        for h, ne in zip(handles, nele):
            e = np.zeros(ne, dtype=np.int64)
            _native.check(lib.xtm_line_copy_entry(h, e.ctypes.data))
            pr = np.zeros(lib.xtm_line_program_size(h), dtype=np.float64)
            _native.check(lib.xtm_line_copy_program(h, pr.ctypes.data))
            entries.append(e)
            progs.append(pr)
    finally:
        _free(lib, handles, b)
    rec_l = np.split(recs, np.cumsum(nele)[:-1])
    # the two ring variants (same cell count, so the same number of bend
    # kicks) have different strengths but the same structures
    fn = [set((r & 0x7ff).tolist()) for r in rec_l]     # function index of each key
    assert fn[0] == fn[1]
    assert len(fn[0]) < 20 and src.count('XT_GEN_FN void rec_') == len(set((recs & 0x7ff).tolist()))
    d = tempfile.mkdtemp(prefix='xtm_batch_')
    cpp, so = os.path.join(d, 'b.cpp'), os.path.join(d, 'b.so')
    open(cpp, 'w').write(src)
    subprocess.run(['g++', '-O1', '-std=c++17', '-fPIC', '-shared', '-ffp-contract=off',
                    '-fno-fast-math', f'-I{ROOT}/include', f'-I{ROOT}/xtrack_amd/csrc', cpp,
                    '-o', so], check=True, capture_output=True)
    glib = ctypes.CDLL(so)
    counts = [300, 40, 200]
    parts = [random_particles(counts[0], seed=1, scale=3e-4, p0c=7e12),
             random_particles(counts[1], seed=2, scale=3e-4, p0c=7e12),
             random_particles(counts[2], seed=3, scale=4e-3, p0c=7e9)]
    for pp in parts:
        pp.q0, pp.mass0, pp.t_sim = 1.0, xt.PROTON_MASS_EV, 0.0
    ref = []
    for f, pp in zip(mk, parts):
        ln = f()
        ln.build_tracker(backend=OracleBackend())
        q = pp.copy()
        ln.track(q, num_turns=3)
        ref.append(q)
    # one host particle set, blocks of at most 256 particles of one line
    starts = np.concatenate([[0], np.cumsum(counts)])
    big = {nn: np.concatenate([pp.to_numpy()[nn] for pp in parts])
           for nn in parts[0].to_numpy()}
    P = xt.Particles.__new__(xt.Particles)
    P._capacity, P._arrays, P._device = int(starts[-1]), big, None
    P.q0, P.mass0, P.t_sim, P.start_tracking_at_element = 1.0, xt.PROTON_MASS_EV, 0.0, -1
    blk = []
    for k in range(3):
        for f in range(starts[k], starts[k + 1], 256):
            blk.append((f, k, min(256, starts[k + 1] - f)))
    blocks = np.array(blk, dtype=[('first', '<i8'), ('line', '<i4'), ('count', '<i4')])
    ptrs = lambda arrs, t: (ctypes.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])
    lengths = np.array([ln.get_length() for ln in lines])
    pst = _native.make_particles_struct(P)
    glib.xtm_batch_host(blocks.ctypes.data_as(ctypes.c_void_p), ctypes.c_int64(len(blocks)),
                        ptrs(rec_l, None), ptrs(entries, None), ptrs(progs, None),
                        nele.ctypes.data_as(ctypes.c_void_p),
                        lengths.ctypes.data_as(ctypes.c_void_p), ctypes.byref(pst),
                        ctypes.c_int32(3), ctypes.c_uint64(0))
    for k in range(3):
        got = xt.Particles.__new__(xt.Particles)
        got._capacity = counts[k]
        got._arrays = {nn: vv[starts[k]:starts[k + 1]].copy() for nn, vv in big.items()}
        got._device = None
        got.q0, got.mass0, got.t_sim, got.start_tracking_at_element = 1.0, xt.PROTON_MASS_EV, 0.0, -1
        bad = compare_exact(got, ref[k])
        assert not bad, (k, bad)
    assert sum(int((r.sorted_by_id()['state'] <= 0).sum()) for r in ref) > 0
