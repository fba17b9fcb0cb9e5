"""The host-side library code (element-program builder, line and batch
code generators, operation counter, the kernel's element code compiled for
the host) and the oracle, built with AddressSanitizer and
UndefinedBehaviorSanitizer (tests/sanitize/) and run over the element
tables of the test lattices: no sanitizer report."""
import os
import subprocess

import numpy as np
import pytest

from helpers import mixed_line, segment_line
from xtrack_amd import lattices
from xtrack_amd.elements import build_raw_table

HERE = os.path.dirname(os.path.abspath(__file__))
SAN = os.path.join(HERE, 'sanitize')


def _write(path, line):
    t, o, p = build_raw_table(line.elements)
    with open(path, 'wb') as fh:
        np.array([len(t), len(p)], dtype=np.int64).tofile(fh)
        t.astype(np.int32).tofile(fh)
        o.astype(np.int64).tofile(fh)
        p.astype(np.float64).tofile(fh)
        np.array([line.get_length()], dtype=np.float64).tofile(fh)


def test_host_code_under_sanitizers(tmp_path):
    r = subprocess.run(['make', '-C', SAN, '-j8'], capture_output=True, text=True)
    if r.returncode != 0:
        pytest.fail(r.stdout + r.stderr)
    from xtrack_amd import Line
    sps = Line.from_json(os.path.join(HERE, 'golden',
                                      'sps_line_no_spacecharge_and_particle.json.gz'))
    lines = {'mixed': mixed_line(n_cells=2), 'segments': segment_line(),
             'errors': lattices.collider_with_errors(0, n_cells=3),
             'thin_fodo': lattices.thin_fodo_ring(), 'sps': sps}
    paths = []
    for name, ln in lines.items():
        pth = str(tmp_path / f'{name}.tab')
        _write(pth, ln)
        paths.append(pth)
    env = dict(os.environ, ASAN_OPTIONS='detect_leaks=1:halt_on_error=1:abort_on_error=0',
               UBSAN_OPTIONS='halt_on_error=1:print_stacktrace=1')
    r = subprocess.run([os.path.join(SAN, '_build', 'sanitize'), *paths], capture_output=True,
                       text=True, env=env, timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert 'runtime error' not in r.stderr and 'AddressSanitizer' not in r.stderr, r.stderr
    assert r.stdout.count(', ok') == len(paths)
