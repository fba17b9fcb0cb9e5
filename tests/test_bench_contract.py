"""The bench's command-line contract on the CPU: one GPU only (a request for
more exits with a message before anything touches a GPU), and the argument
defaults of the headline run."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, 'bench.py'), '--gpus', '2'],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert 'one GPU' in r.stderr and r.stdout == ''


def test_headline_defaults():
    sys.path.insert(0, ROOT)
    import importlib
    bench = importlib.import_module('bench')
    old = sys.argv
    try:
        sys.argv = ['bench.py']
        a = bench.parse()
    finally:
        sys.argv = old
    assert (a.gpus, a.steps, a.warmup, a.particles, a.turns_per_step) == (1, 10, 1, 1_000_000, 1)
    assert a.arithmetic == 'exact' and a.workload == 'collider' and a.traffic == 'measure'
    assert not a.batch_interpreter
