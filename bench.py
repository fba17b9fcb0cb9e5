"""Benchmark: multi-turn FP64 tracking throughput (particle-turns/s) on one
MI355X for the thick collider-scale ring (28,338 elements: 1104 thick
Yoshida-4 sector bends, 368 thick quadrupoles, 368 sextupoles matched to
chromaticity +2, 184 thin octupoles, drifts and a LimitRect after every
element, one 16 MV 400 MHz cavity) with 1,000,000 7 TeV protons, Gaussian
6D beam matched to the ring (normalised emittances 3.75 um, sigma_z 0.0755
m, sigma_delta 1.13e-4).

A step tracks every particle through `--turns-per-step` turns (one kernel
launch); inputs are resident in HBM before the timed region.  Prints one
JSON line.  Also reports the roofline of the tracking kernel (algorithmic
FP64 operations counted by running the kernel's element code on the host
with an operation-counting scalar, over the kernel time measured with HIP
events on the launch stream) and a CPU baseline: the CPU restatement of
the reference tracker (oracle/, platform libm, OpenMP with the reference's
chunking) timed on a bounded sample of the same workload on this host.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6
# sextupole strengths for chromaticity (+2, +2) of the ring below, obtained
# with xtrack_amd.lattices.match_chromaticity; re-derived when --rematch
K2F, K2D = 0.0319663, -0.06119144
NEMITT = 3.75e-6
SIGMA_Z = 0.0755
SIGMA_DELTA = 1.13e-4
FLOP_PER_CALL = 20     # weight of one elementary-function call


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=5)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--turns-per-step', type=int, default=2)
    ap.add_argument('--particles', type=int, default=1_000_000)
    ap.add_argument('--cpu-sample-particles', type=int, default=2048)
    ap.add_argument('--cpu-sample-turns', type=int, default=10)
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--rematch', action='store_true')
    return ap.parse_args()


def cpu_baseline(line_fn, tw, n, turns):
    """Oracle (platform libm) on `n` particles x `turns` turns with all host
    threads the process may use."""
    sys.path.insert(0, os.path.join(ROOT, 'tests'))
    from oracle_backend import OracleBackend
    from xtrack_amd import optics
    threads = len(os.sched_getaffinity(0))
    line = line_fn()
    line.build_tracker(backend=OracleBackend(math='libm', num_threads=threads))
    p = optics.matched_gaussian(line, line.particle_ref, n, NEMITT, NEMITT, SIGMA_Z,
                                SIGMA_DELTA, seed=2, tw=tw)
    t0 = time.perf_counter()
    line.track(p, num_turns=turns)
    dt = time.perf_counter() - t0
    done = int(np.sum(p.to_numpy()['at_turn']))
    return {'value': done / dt, 'unit': 'particle-turns/s', 'cores': threads,
            'kind': 'port',
            'sample': f'{n} particles x {turns} turns of the same ring and beam '
                      f'({done} particle-turns in {dt:.1f} s), CPU restatement of the '
                      f'reference tracker with platform libm, OpenMP {threads} threads'}


def main():
    args = parse()
    if args.gpus != 1:
        print(f'bench.py: this build measures one GPU only (asked for {args.gpus}); '
              'particles shard across GPUs with no exchange during tracking, '
              'so N GPUs give N times the per-GPU rate.', file=sys.stderr)
        sys.exit(2)
    import torch
    from xtrack_amd import lattices, optics, _native

    k2f, k2d = K2F, K2D
    if args.rematch:
        (k2f, k2d), _ = lattices.match_chromaticity(
            lambda a, b: lattices.collider_ring(k2f=a, k2d=b))

    def make_line():
        return lattices.collider_ring(k2f=k2f, k2d=k2d)

    line = make_line()
    tracker = line.build_tracker(device='cuda:0')
    ref = line.particle_ref
    tw = optics.twiss(line, ref)
    chroma = optics.chromaticity(line, ref)
    parts = optics.matched_gaussian(line, ref, args.particles, NEMITT, NEMITT, SIGMA_Z,
                                    SIGMA_DELTA, seed=1, tw=tw)
    parts.move(_context='cuda:0')

    # algorithmic operations per particle-turn (kernel element code, counted)
    lib = _native.load()
    ops = np.zeros(6)
    coords = np.array([1e-5, 1e-7, 1e-5, 1e-7, 1e-3, 1e-5])
    _native.check(lib.xtm_line_count_ops(tracker._handle.ptr, coords.ctypes.data,
                                         float(ref._np('p0c')[0]), ref.mass0, ref.q0,
                                         ops.ctypes.data))
    flop_pt = ops[5] + FLOP_PER_CALL * ops[4]

    tps = args.turns_per_step
    for _ in range(args.warmup):
        line.track(parts, num_turns=tps)
    torch.cuda.synchronize()
    at_turn0 = parts.at_turn.sum().item()
    stream = torch.cuda.current_stream()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        ev[k][0].record(stream)
        line.track(parts, num_turns=tps)
        ev[k][1].record(stream)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    at_turn1 = parts.at_turn.sum().item()
    pturns = at_turn1 - at_turn0
    kern_ms = [a.elapsed_time(b) for a, b in ev]
    n_alive = int((parts.state > 0).sum().item())

    value = pturns / elapsed
    # per launch: all particles alive at the start of the step x tps turns
    kern_s = float(np.mean(kern_ms)) / 1e3
    achieved = flop_pt * (pturns / args.steps) / kern_s / 1e12
    out = {
        'metric': 'particle-turns/s (FP64) on 1 MI355X',
        'value': value,
        'unit': 'particle-turns/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': elapsed / args.steps * 1e3,
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f64',
        'data': 'synthetic (seeded matched Gaussian beam, synthetic lattice)',
        'config': {'workload': 'collider ring, 8 arcs x 23 FODO cells, '
                               f'{len(line.elements)} elements, 7 TeV protons',
                   'particles': args.particles, 'turns_per_step': tps,
                   'particles_alive_end': n_alive,
                   'tunes': [round(float(tw['qx']), 4), round(float(tw['qy']), 4)],
                   'chromaticity': [round(float(chroma[0]), 3), round(float(chroma[1]), 3)]},
        'roofline': {'bound': 'fp64-valu', 'achieved': achieved, 'peak': PEAK_FP64_TFLOPS,
                     'unit': 'TFLOP/s', 'frac': achieved / PEAK_FP64_TFLOPS, 'traffic': None,
                     'flop_per_particle_turn': flop_pt,
                     'ops_per_particle_turn': {'add': ops[0], 'mul': ops[1], 'div': ops[2],
                                               'sqrt': ops[3], 'calls': ops[4]},
                     'kernel_ms_mean': float(np.mean(kern_ms))},
    }
    if not args.no_cpu_baseline:
        out['cpu_baseline'] = cpu_baseline(make_line, tw, args.cpu_sample_particles,
                                           args.cpu_sample_turns)
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
