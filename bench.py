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

HBM traffic of the tracking kernel is measured by running this script twice
more under ``rocprofv3 --pmc`` (FETCH_SIZE, then WRITE_SIZE) as child
processes before this process touches the GPU; each child also runs a plain
copy of known size, whose counter readings convert the counters to bytes.
"""
import argparse
import csv
import glob
import json
import os
import shutil
import subprocess
import sys
import tempfile
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
# particle state moved per particle and launch: 13 fields of 8 bytes read
# (x px y py zeta delta rpp rvv s chi at_element at_turn state), 12 written
# back (all but chi); the cavity and bend edges read/write a few more fields
# in place (ptau beta0 charge_ratio p0c ax ay), counted by the traffic runs
STATE_BYTES_PER_PARTICLE = (13 + 12) * 8
COPY_MIB = 256
_T0 = time.perf_counter()


def log(msg):
    """Progress line on stderr (each stage of a run prints one)."""
    print(f'bench.py [{time.perf_counter() - _T0:7.1f} s]: {msg}', file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=10)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--turns-per-step', type=int, default=1)
    ap.add_argument('--particles', type=int, default=1_000_000)
    ap.add_argument('--cpu-sample-particles', type=int, default=8192)
    ap.add_argument('--cpu-sample-turns', type=int, default=24)
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--arithmetic', choices=('exact', 'fast'), default='exact',
                    help='arithmetic of the headline run: exact (IEEE quotients in the '
                         "reference's order) or fast (re-associated polar-drift path length)")
    ap.add_argument('--no-variants', action='store_true',
                    help='skip the timed run of the other arithmetic')
    ap.add_argument('--no-batch', action='store_true',
                    help='skip the batched-variants leg of the default run')
    ap.add_argument('--traffic', choices=('measure', 'none'), default='measure',
                    help='measure HBM traffic with two rocprofv3 --pmc child runs')
    ap.add_argument('--copy-calibration', action='store_true',
                    help='run a plain device copy of known size before tracking '
                         '(used by the traffic child runs)')
    ap.add_argument('--compile', choices=('auto', 'on', 'off'), default='auto',
                    help='track with compiled line code (default when possible) or the '
                         'element-program interpreter')
    ap.add_argument('--waves-per-eu', type=int, default=0,
                    help='occupancy target of the compiled kernel (0: library default)')
    ap.add_argument('--rematch', action='store_true')
    ap.add_argument('--ring-error-seed', type=int, default=None,
                    help='track the ring with the b3..b6 error set of this seed on every '
                         'magnet (lattices.collider_with_errors): structure code with '
                         'constants loaded from the element programs, for profiling')
    ap.add_argument('--ring-error-scale', type=float, default=1e-4)
    ap.add_argument('--da-angles', type=int, default=45)
    ap.add_argument('--da-amplitudes', type=int, default=400)
    ap.add_argument('--da-turns', type=int, default=2000)
    ap.add_argument('--da-seeds', type=int, default=5)
    ap.add_argument('--da-feature-turns', type=int, default=64)
    ap.add_argument('--da-label-fraction', type=float, default=0.03)
    ap.add_argument('--da-band-sigma', type=float, default=0.3)
    ap.add_argument('--da-grid-sizes', default=None,
                    help='lo,hi: particles per seed drawn log-uniformly in [lo, hi] (seeded), '
                         'amplitude count per seed = particles / (angles x 2); default: '
                         '--da-amplitudes for every seed')
    ap.add_argument('--da-epochs', type=int, default=100)
    ap.add_argument('--da-error-scale', type=float, default=3e-3,
                    help='relative strength of the b3..b6 errors (at 1e-4 the aperture of '
                         'the ring, 57-77 sigma, bounds the 0-20 sigma grid from outside)')
    ap.add_argument('--da-predictors', default='mlp',
                    help='comma-separated: mlp (the network), interp (the same screening '
                         'with per-angle interpolation of the labels instead)')
    ap.add_argument('--workload', choices=('collider', 'batch', 'da', 'mlp'), default='collider',
                    help='collider: the thick ring (the headline line); batch: 64 lattice '
                         'variants of it with uneven particle counts in one launch; da: '
                         'dynamic-aperture scan of error seeds of the ring, plain against '
                         'screened by the surrogate')
    ap.add_argument('--batch-variants', type=int, default=64)
    ap.add_argument('--batch-max-particles', type=int, default=200_000)
    ap.add_argument('--batch-graph', action='store_true',
                    help='replay each step from a captured HIP graph')
    ap.add_argument('--batch-compile', action='store_true',
                    help='(the default) compile the variants into one kernel (element '
                         'structures with constants read from each line)')
    ap.add_argument('--batch-interpreter', action='store_true',
                    help='track the batch with the element-program interpreter instead')
    ap.add_argument('--batch-waves-per-eu', type=int, default=0,
                    help='occupancy target of the compiled batch code (0: library default)')
    return ap.parse_args()


def _read_counter(outdir, counter):
    """(copy_value, track_value) of `counter` from a rocprofv3 csv run: the
    copy is the calibration dispatch (kernel name holds 'neg'), the tracking
    value is that of the last dispatch (the timed launch) of the kernel that
    tracked the beam: the compiled line code (xtm_line_kernel) when the line
    is compiled, else the interpreter (track_kernel).  With compiled code the
    interpreter kernel is still launched after each compiled launch, for the
    particles with chi != 1 (none here), and is not the one measured."""
    files = glob.glob(os.path.join(outdir, '**', '*counter_collection.csv'), recursive=True)
    if not files:
        raise RuntimeError(f'no counter_collection.csv under {outdir}')
    per_dispatch = {}
    for fn in files:
        with open(fn, newline='') as fh:
            for row in csv.DictReader(fh):
                if row.get('Counter_Name') != counter:
                    continue
                key = (fn, int(row['Dispatch_Id']))
                name = row['Kernel_Name']
                val = float(row['Counter_Value'])
                old = per_dispatch.get(key, (name, 0.0))
                per_dispatch[key] = (name, old[1] + val)
    items = sorted(per_dispatch.items(), key=lambda kv: kv[0][1])
    copy = [v for _, (nm, v) in items if 'neg' in nm]
    track = [v for _, (nm, v) in items if 'xtm_line_kernel' in nm]
    if not track:
        track = [v for _, (nm, v) in items if 'track_kernel' in nm]
    if not copy or not track:
        raise RuntimeError(f'{counter}: calibration or tracking dispatch not found')
    return max(copy), track[-1]


def measure_traffic(args):
    """HBM bytes per tracking launch from two rocprofv3 --pmc child runs
    (FETCH_SIZE, then WRITE_SIZE, in KiB per dispatch), each scaled by the
    reading of the same counter for a copy of COPY_MIB MiB (reads and writes
    exactly that many bytes)."""
    prof = shutil.which('rocprofv3')
    if prof is None:
        raise RuntimeError('rocprofv3 not on PATH')
    copy_bytes = COPY_MIB * 2 ** 20
    res = {}
    tmp = tempfile.mkdtemp(prefix='xtm_traffic_')
    try:
        for counter in ('FETCH_SIZE', 'WRITE_SIZE'):
            outdir = os.path.join(tmp, counter)
            cmd = [prof, '--pmc', counter, '--output-format', 'csv', '-d', outdir, '-o', 'run',
                   '--', sys.executable, os.path.abspath(__file__), '--steps', '1',
                   '--warmup', '0', '--turns-per-step', str(args.turns_per_step),
                   '--particles', str(args.particles), '--no-cpu-baseline',
                   '--traffic', 'none', '--copy-calibration', '--compile', args.compile,
                   '--arithmetic', args.arithmetic, '--no-variants', '--no-batch',
                   '--waves-per-eu', str(args.waves_per_eu)]
            log(f'traffic pass {counter}')
            with open(os.path.join(tmp, counter + '.log'), 'w') as fh:
                subprocess.run(cmd, stdout=fh, stderr=subprocess.STDOUT, timeout=420,
                               check=True)
            kib_copy, kib_track = _read_counter(outdir, counter)
            scale = copy_bytes / (kib_copy * 1024.0)
            res[counter] = {'kib_track': kib_track, 'kib_copy': kib_copy, 'scale': scale,
                            'bytes': kib_track * 1024.0 * scale}
            log(f'{counter} track {kib_track:.0f} KiB, copy {kib_copy:.0f} KiB '
                f'(scale {scale:.4f})')
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    return res


def _oracle_backend(math, threads):
    sys.path.insert(0, os.path.join(ROOT, 'tests'))
    from oracle_backend import OracleBackend
    return OracleBackend(math=math, num_threads=threads)


def cpu_leg(line_fn, parts, turns, threads, what):
    """Oracle (platform libm: the reference's CPU arithmetic and its OpenMP
    chunking) tracking `parts` for `turns` turns on `threads` host threads;
    particle-turns/s."""
    line = line_fn()
    line.build_tracker(backend=_oracle_backend('libm', threads))
    p = parts.copy()
    t0 = time.perf_counter()
    line.track(p, num_turns=turns)
    dt = time.perf_counter() - t0
    done = int(np.sum(p.to_numpy()['at_turn']) - np.sum(parts.to_numpy()['at_turn']))
    return {'value': done / dt, 'unit': 'particle-turns/s', 'cores': threads,
            'kind': 'port',
            'sample': f'{what}: {p._capacity} particles x {turns} turns ({done} particle-turns '
                      f'in {dt:.1f} s), CPU restatement of the reference tracker with '
                      f'platform libm, OpenMP {threads} thread(s)'}


def cpu_baseline(make_line, beam, thin_line, thin_beam, args):
    """The reported CPU baseline: the collider ring on a bounded sample of
    its beam with every host thread, plus a one-thread leg of the same ring
    and the thin FODO workload at its full size (10,000 particles x 100
    turns) with every thread and with one."""
    threads = len(os.sched_getaffinity(0))
    main_leg = cpu_leg(make_line, _subset(beam, np.arange(args.cpu_sample_particles)),
                       args.cpu_sample_turns, threads, 'collider ring, matched beam')
    main_leg['legs'] = {
        'collider_1_thread': cpu_leg(make_line, _subset(beam, np.arange(512)), 4, 1,
                                     'collider ring, matched beam'),
        'thin_fodo_full': cpu_leg(thin_line, thin_beam, 100, threads,
                                  'thin FODO ring (10,000 particles, 100 turns)'),
        'thin_fodo_full_1_thread': cpu_leg(thin_line, thin_beam, 100, 1,
                                           'thin FODO ring (10,000 particles, 100 turns)'),
    }
    return main_leg


def _subset(parts, idx):
    """Particles holding entries `idx` of the host set `parts` (same
    particle ids)."""
    import xtrack_amd as xt
    h = parts.to_numpy()
    out = xt.Particles(p0c=float(h['p0c'][0]), mass0=parts.mass0, q0=parts.q0,
                       x=np.zeros(len(idx)))
    out._set_all({nn: np.ascontiguousarray(vv[idx]) for nn, vv in h.items()})
    return out


COORDS = ('x', 'px', 'y', 'py', 'zeta', 'delta')


def _compare(got, ref):
    """Loss records and coordinate differences of two sorted particle
    dicts: losses_equal (state, at_turn, at_element identical), the number
    lost, identical bits of every compared field, and per coordinate
    max |got - ref| / max |ref| over the particles alive in both."""
    sys.path.insert(0, os.path.join(ROOT, 'tests'))
    from helpers import COMPARED
    same = all(np.array_equal(got[ff], ref[ff]) for ff in ('state', 'at_turn', 'at_element'))
    bits = True
    for ff in COMPARED:
        a, b = got[ff], ref[ff]
        if a.dtype.kind == 'f':
            eq = ((a == b) & (np.signbit(a) == np.signbit(b))) | (np.isnan(a) & np.isnan(b))
        else:
            eq = a == b
        bits = bits and bool(eq.all())
    alive = (got['state'] > 0) & (ref['state'] > 0)
    err = {}
    for ff in COORDS:
        den = float(np.max(np.abs(ref[ff][alive]))) if alive.any() else 0.0
        num = float(np.max(np.abs(got[ff][alive] - ref[ff][alive]))) if alive.any() else 0.0
        err[ff] = num / den if den > 0 else num
    return {'losses_equal': bool(same), 'lost': int((ref['state'] <= 0).sum()),
            'bit_identical': bits, 'max_rel_err': max(err.values()), 'rel_err': err}


def _oracle_track(line_fn, parts, turns, math, threads, snapshots=(), log_every=None,
                  what=''):
    """Oracle track of `parts`; sorted copies after each turn count in
    `snapshots` and at the end; a progress line every `log_every` turns."""
    line = line_fn()
    line.build_tracker(backend=_oracle_backend(math, threads))
    p = parts.copy()
    out, done = {}, 0
    stops = set(snapshots) | {turns}
    if log_every:
        stops |= set(range(log_every, turns, log_every))
    for t in sorted(stops):
        line.track(p, num_turns=t - done)
        done = t
        if t in snapshots or t == turns:
            out[t] = p.sorted_by_id()
        if log_every:
            log(f'oracle ({math}): {what}, turn {t} of {turns}')
    return out


def collider_parity_sample(beam, n_core=48, n_edge=16, seed=3):
    """Particles of the collider beam plus particles at the edge of the
    horizontal aperture (amplitude 7.5-8.7 mm at the start, where the
    dispersion is 2.5 m), which synchrotron motion loses at scattered
    turns."""
    import xtrack_amd as xt
    core = _subset(beam, np.arange(n_core)).to_numpy()
    rng = np.random.default_rng(seed)
    edge = dict(x=rng.uniform(0.0075, 0.0087, n_edge) * rng.choice([-1, 1], n_edge),
                px=rng.normal(0, 3e-6, n_edge), y=rng.normal(0, 1e-3, n_edge),
                py=rng.normal(0, 3e-6, n_edge), zeta=rng.normal(0, 0.1, n_edge),
                delta=rng.normal(0, 2e-4, n_edge))
    kw = {ff: np.concatenate([core[ff][:n_core], edge[ff]]) for ff in edge}
    return xt.Particles(p0c=float(core['p0c'][0]), mass0=beam.mass0, q0=beam.q0, **kw)


def parity_and_cpu_legs(args, make_line, parts_gpu, init, turns_done):
    """GPU against the oracle on the same inputs, after the timed region:
    (1) a sample of the timed beam after the bench's turns; (2) the thin
    FODO workload at its full size (10,000 matched particles) after 1 and
    1000 turns; (3) 64 particles of the collider workload (48 of the
    matched beam, 16 at the aperture edge) after 1 and 1000 turns in each
    arithmetic, tracked on two more streams while the CPU runs the oracle
    and the baseline legs.  Each against the oracle with the project's math
    library (the kernel's elementary functions: in exact arithmetic the bits
    must be identical) and with the platform libm (the reference's
    arithmetic: losses identical, coordinate error stated)."""
    import torch
    from xtrack_amd import lattices, optics
    threads = len(os.sched_getaffinity(0))
    res = {'tolerance': {'after_1_turn': 1e-12, 'after_1000_turns': 1e-9},
           'error_definition': 'per coordinate max |GPU - oracle| / max |oracle| over '
                               'particles alive in both; losses_equal: state, at_turn and '
                               'at_element identical'}
    log(f'parity legs ({threads} host threads)')
    # (1) sample of the timed beam
    rng = np.random.default_rng(77)
    idx = np.sort(rng.choice(init._capacity, 256, replace=False))
    g_beam = _subset(parts_gpu, idx).sorted_by_id()
    samp = _subset(init, idx)
    # (2) thin FODO workload on the GPU
    thin = lattices.thin_fodo_ring()
    thin.build_tracker(device='cuda:0')
    thin_beam = optics.matched_gaussian(thin, thin.particle_ref, 10_000, 2.5e-6, 2.5e-6, 0.2,
                                        1e-3, seed=0)
    pg = thin_beam.copy(_context='cuda:0')
    thin.track(pg, num_turns=1)
    g_thin = {1: pg.sorted_by_id()}
    thin.track(pg, num_turns=999)
    g_thin[1000] = pg.sorted_by_id()
    log('thin FODO workload tracked on the GPU (1000 turns)')
    # (3) collider sample in both arithmetics: 1 turn, then 999 turns on
    # their own streams
    csamp = collider_parity_sample(init)
    g_col, runs = {}, []
    for arith in ('exact', 'fast'):
        cline = make_line()
        cline.build_tracker(device='cuda:0', arithmetic=arith)
        cg = csamp.copy(_context='cuda:0')
        st = torch.cuda.Stream()
        torch.cuda.synchronize()
        with torch.cuda.stream(st):
            cline.track(cg, num_turns=1)
            st.synchronize()
            g_col[arith] = {1: cg.sorted_by_id()}
            cline.track(cg, num_turns=999)
        runs.append((arith, cline, cg, st))
    log('collider sample: 1 turn done, 999 turns queued')
    # CPU work while the GPU tracks
    o_beam = {m: _oracle_track(make_line, samp, turns_done, m, threads)[turns_done]
              for m in ('xm', 'libm')}
    log('oracle: beam sample')
    o_thin = {m: _oracle_track(lattices.thin_fodo_ring, thin_beam, 1000, m, threads, (1,))
              for m in ('xm', 'libm')}
    log('oracle: thin FODO workload')
    # 64 particles: one host thread is faster than the OpenMP chunking of
    # the restated CPU tracker over 16 threads (synchronisation per element)
    o_col = {m: _oracle_track(make_line, csamp, 1000, m, 1, (1,), log_every=250,
                              what='collider sample') for m in ('xm', 'libm')}
    log('oracle: collider sample')
    cpu = cpu_baseline(make_line, init, lattices.thin_fodo_ring, thin_beam, args)
    log('CPU baseline legs')
    for arith, cline, cg, st in runs:
        st.synchronize()
        with torch.cuda.stream(st):
            g_col[arith][1000] = cg.sorted_by_id()
    log('collider sample: 1000 turns done on the GPU')

    def both(g, o):
        return {'vs_oracle_project_math': _compare(g, o['xm']),
                'vs_oracle_libm': _compare(g, o['libm'])}

    res['collider_beam_sample'] = dict(particles=len(idx), turns=turns_done,
                                       arithmetic=args.arithmetic, **both(g_beam, o_beam))
    res['thin_fodo_10k'] = {'particles': 10_000,
                            'after_1_turn': both(g_thin[1], {m: o_thin[m][1] for m in o_thin}),
                            'after_1000_turns': both(g_thin[1000],
                                                     {m: o_thin[m][1000] for m in o_thin})}
    for arith in ('exact', 'fast'):
        res[f'collider_64_{arith}'] = {
            'particles': 64, 'of_which_at_aperture_edge': 16, 'arithmetic': arith,
            'after_1_turn': both(g_col[arith][1], {m: o_col[m][1] for m in o_col}),
            'after_1000_turns': both(g_col[arith][1000], {m: o_col[m][1000] for m in o_col})}
    return res, cpu


def batch_leg(args, steps, kernels=None):
    """Batched lattice variants: `--batch-variants` variants of the ring
    (seeds 0.., cells per arc uniform in 8..40, quadrupole and sextupole
    strengths scattered by N(0, 1e-4) relative), particle counts log-uniform
    in [1000, --batch-max-particles], all tracked in one launch per step of
    `--turns-per-step` turns.  Beams: Gaussian, the same small emittance for
    every variant (not matched to each variant's optics).  Each kernel of
    `kernels` ('compiled': generated batch code, 'interpreter': element
    programs) is autotuned and timed on its own copy of the beams; the
    measurement of the faster one is returned, the other's time beside it."""
    import torch
    import xtrack_amd as xt
    from xtrack_amd import lattices, _native
    if kernels is None:
        kernels = ('interpreter',) if args.batch_interpreter else ('compiled',)
    rng = np.random.default_rng(12345)
    nv = args.batch_variants
    counts = np.exp(rng.uniform(np.log(1000), np.log(args.batch_max_particles), nv))
    counts = counts.astype(np.int64)
    t0 = time.perf_counter()
    lines = [lattices.collider_variant(s, k2f=K2F, k2d=K2D) for s in range(nv)]
    beams = []
    for k, n in enumerate(counts):
        r = np.random.default_rng(1000 + k)
        beams.append(xt.Particles(p0c=7e12, mass0=xt.PROTON_MASS_EV,
                                  x=r.normal(0, 3e-4, n), px=r.normal(0, 3e-6, n),
                                  y=r.normal(0, 3e-4, n), py=r.normal(0, 3e-6, n),
                                  zeta=r.normal(0, SIGMA_Z, n),
                                  delta=r.normal(0, SIGMA_DELTA, n)))
    log(f'batch: {nv} lines built ({time.perf_counter() - t0:.1f} s)')
    nele = [len(ln.elements) for ln in lines]
    res = {}
    flop = None
    for kern in kernels:
        t1 = time.perf_counter()
        bt = xt.BatchTracker(lines, device='cuda:0', compile=(kern == 'compiled'),
                             waves_per_eu=args.batch_waves_per_eu)
        setup_s = time.perf_counter() - t1
        log(f'batch ({kern}): tracker ready ({setup_s:.1f} s)')
        if flop is None:
            # operations per particle-turn of each variant (same counting as the ring)
            lib = _native.load()
            flop = np.zeros(nv)
            coords = np.array([1e-5, 1e-7, 1e-5, 1e-7, 1e-3, 1e-5])
            for k, h in enumerate(bt._handles):
                ops = np.zeros(6)
                _native.check(lib.xtm_line_count_ops(h.ptr, coords.ctypes.data, 7e12,
                                                     xt.PROTON_MASS_EV, 1.0, ops.ctypes.data))
                flop[k] = ops[5] + FLOP_PER_CALL * ops[4]
        parts = [b.copy() for b in beams]
        cfg = bt.tune(parts, turns=1)
        log(f'batch ({kern}): autotuned {cfg}')
        tps = args.turns_per_step
        for _ in range(args.warmup):
            bt.track(parts, num_turns=tps, turns_per_launch=tps, use_graph=args.batch_graph)
        before = np.array([int(p.to_numpy()['at_turn'].sum()) for p in parts])
        alive0 = np.array([int((p.to_numpy()['state'] > 0).sum()) for p in parts])
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        info = bt.track(parts, num_turns=tps * steps, turns_per_launch=tps,
                        use_graph=args.batch_graph)
        elapsed = time.perf_counter() - t1
        after = np.array([int(p.to_numpy()['at_turn'].sum()) for p in parts])
        pturns = int((after - before).sum())
        kern_s = info['kernel_ms'] / 1e3
        achieved = float(np.sum(flop * (after - before))) / kern_s / 1e12
        log(f'batch ({kern}): {steps} steps, {info["kernel_ms"] / steps:.1f} ms per step')
        res[kern] = {
            'value': pturns / kern_s, 'unit': 'particle-turns/s',
            'ms_per_step': info['kernel_ms'] / steps, 'steps': steps,
            'workload': f'{nv} variants of the collider ring (cells per arc 8..40, '
                        'quad/sextupole errors 1e-4), one launch per step',
            'variants': nv, 'elements_min_max': [min(nele), max(nele)],
            'particles_total': int(counts.sum()),
            'particles_min_max': [int(counts.min()), int(counts.max())],
            'turns_per_step': tps,
            'kernel': ('compiled batch code (literal functions for shared records, '
                       'structure functions for the rest)' if bt.is_compiled
                       else 'element-program interpreter (batch)'),
            'graph': bool(info['graph']), 'autotuned': list(cfg),
            'tune_ms': {f'{k[0]}x{k[1]}': round(v, 2) for k, v in bt.tune_timings.items()},
            'alive_start_end': [int(alive0.sum()),
                                int(sum((p.to_numpy()['state'] > 0).sum() for p in parts))],
            'host_setup_s': round(setup_s, 1), 'wall_s': round(elapsed, 2),
            'roofline': {'bound': 'fp64-valu', 'achieved': achieved,
                         'peak': PEAK_FP64_TFLOPS, 'unit': 'TFLOP/s',
                         'frac': achieved / PEAK_FP64_TFLOPS},
        }
        del bt, parts
    best = max(res, key=lambda kk: res[kk]['value'])
    out = dict(res[best])
    out['other_kernels_ms_per_step'] = {kk: round(v['ms_per_step'], 2)
                                        for kk, v in res.items() if kk != best}
    return out


def main_mlp(args):
    """The survival network's products on the matrix cores (xtm_gemm_bf16,
    BF16 operands, FP32 accumulation, fused bias, ReLU and BF16 rounding of
    the output): the hidden-layer product of inference over a DA grid,
    (65,536 x 256) x (256 x 256), timed with HIP events on the launch
    stream over `--steps` launches, and the training step of the
    16-256-256-256-1 network at the minibatch of the fit (1,024 samples:
    forward, backward and Adam), launched kernel by kernel and replayed from
    a captured HIP graph (wall time per step over 100 steps).  The product
    moves 32 MiB of BF16 input,
    64 MiB of FP32 and 32 MiB of BF16 output per launch for 8.6 GFLOP
    (64 flop per byte, below the 312 flop per byte where MI355X's BF16
    matrix peak and its HBM bandwidth meet), so it is bound by HBM."""
    import torch
    from xtrack_amd import surrogate as S
    dev = torch.device('cuda:0')
    g = torch.Generator(device='cpu').manual_seed(0)
    M, K, N = 65536, 256, 256
    a = S.to_bf16(torch.randn(M, K, generator=g).to(dev))
    w = S.to_bf16((torch.randn(K, N, generator=g) / 16).to(dev))
    bias = torch.zeros(N, dtype=torch.float32, device=dev)
    for _ in range(max(args.warmup, 1)):
        S.gemm(a, w, bias=bias, relu=True, want_bf16=True)
    stream = torch.cuda.current_stream()
    steps = max(args.steps, 1)
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        ev[k][0].record(stream)
        S.gemm(a, w, bias=bias, relu=True, want_bf16=True)
        ev[k][1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = [x.elapsed_time(y) for x, y in ev]
    ms_mean = float(np.mean(ms))
    flop = 2.0 * M * N * K
    bytes_ = 2.0 * M * K + 2.0 * K * N + 4.0 * N + (4.0 + 2.0) * M * N
    # training steps at the fit's minibatch (1,024 of 4,096 samples), launched
    # one by one and replayed from the captured graph (one eager step and
    # the capture included, over 100 steps)
    xb = torch.randn(4096, 16, generator=g).to(dev)
    yb = torch.randn(4096, generator=g).to(dev)
    train_ms = {}
    for mode in ('eager', 'graph'):
        net = S.MLP(device=dev, seed=0)
        net.fit(xb, yb, epochs=1, batch=1024, seed=0, graph=(mode == 'graph'))
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        net.fit(xb, yb, epochs=25, batch=1024, seed=0, graph=(mode == 'graph'))
        torch.cuda.synchronize()
        train_ms[mode] = round((time.perf_counter() - t1) / 100 * 1e3, 4)
    achieved = bytes_ / (ms_mean * 1e-3) / 1e9
    out = {
        'metric': 'survival-network hidden-layer product, HBM GB/s', 'value': achieved,
        'unit': 'GB/s', 'n_gpus': 1, 'steps': steps, 'warmup': args.warmup,
        'ms_per_step': wall / steps * 1e3, 'higher_is_better': True, 'scaling': 'weak',
        'vs_baseline': None, 'dtype': 'bf16 operands, f32 accumulate',
        'data': 'synthetic (Gaussian operands)',
        'config': {'workload': f'({M} x {K}) x ({K} x {N}) product with bias, ReLU and BF16 '
                               f'output (inference of the 16-256-256-256-1 network over '
                               f'{M} particles, hidden layer)',
                   'kernel_ms': [round(v, 4) for v in ms], 'tflops': flop / (ms_mean * 1e-3) / 1e12,
                   'training_step_ms_batch_1024': train_ms},
        'roofline': {'bound': 'hbm', 'achieved': achieved, 'peak': 8000.0, 'unit': 'GB/s',
                     'frac': achieved / 8000.0, 'traffic': None,
                     'algorithmic_bytes_per_launch': bytes_, 'flop_per_launch': flop,
                     'mfma_frac': flop / (ms_mean * 1e-3) / 2.5e15},
    }
    print(json.dumps(out), flush=True)


def main_batch(args):
    """`--workload batch`: the batched-variants measurement as the line."""
    b = batch_leg(args, args.steps)
    out = {
        'metric': 'particle-turns/s (FP64) on 1 MI355X, batched lattice variants',
        'value': b['value'], 'unit': 'particle-turns/s', 'n_gpus': 1,
        'steps': args.steps, 'warmup': args.warmup, 'ms_per_step': b['ms_per_step'],
        'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f64',
        'data': 'synthetic (seeded Gaussian beams, synthetic lattice variants)',
        'config': {k: v for k, v in b.items()
                   if k not in ('value', 'unit', 'ms_per_step', 'steps', 'roofline')},
        'roofline': dict(b['roofline'], compute_pipe='FP64 VALU (MI355X FP64 peak 78.6 '
                                                     'TFLOP/s; no matrix instructions '
                                                     'issued)', traffic=None),
    }
    print(json.dumps(out), flush=True)


def main_da(args):
    """Dynamic-aperture scan of the collider ring with --da-seeds seeded
    b3..b6 error sets (all seeds in one batch), plain against screened by
    the survival surrogate and against the same screening without the
    network (per-angle interpolation of the labels): polar grid of
    --da-angles angles x --da-amplitudes amplitudes (0-20 sigma, normalised
    emittance 3.75 um) x delta in {0, 2.7e-4} per seed, --da-turns turns.
    The errors are N(0, --da-error-scale) relative to the main field at
    17 mm.
    Reports the wall times of the three scans (grid built and resident on
    the GPU before timing), the fraction of the grid each tracks in full
    and the aperture error per seed against the plain scan."""
    import torch
    import xtrack_amd as xt
    from xtrack_amd import lattices, surrogate
    t0 = time.perf_counter()
    mk = lambda s_: lattices.collider_with_errors(s_, k2f=K2F, k2d=K2D,
                                                  rel=args.da_error_scale)
    lines = [mk(s_) for s_ in range(args.da_seeds)]
    for ln in lines:
        ln.build_tracker(device='cuda:0')
    bt = xt.BatchTracker(lines, device='cuda:0', compile=True)
    log(f'da: {len(lines)} error-seed lattices compiled ({time.perf_counter() - t0:.1f} s)')
    if args.da_grid_sizes:
        lo, hi = (float(v) for v in args.da_grid_sizes.split(','))
        rng = np.random.default_rng(777)
        sizes = np.exp(rng.uniform(np.log(lo), np.log(hi), len(lines)))
        n_r = [max(2, int(round(v / (2 * args.da_angles)))) for v in sizes]
    else:
        n_r = args.da_amplitudes
    gs = surrogate.GridSet(lines, 20.0, n_r, args.da_angles, (0.0, 2.7e-4),
                           NEMITT, NEMITT, torch.device('cuda:0'))
    norm = surrogate.normal_form_4d(lines[0], torch.device('cuda:0'))
    n = gs.n_slots
    log(f'da: grid of {n} particles built ({time.perf_counter() - t0:.1f} s)')
    T = args.da_turns
    plain = surrogate.plain_da_scan_batch(
        bt, gs, T, progress=lambda d, n_: log(f'da: plain, turn {d}, {n_} tracked'))
    log(f'da: plain scan {plain["time"]:.1f} s')
    out_scans = {}
    for pred in args.da_predictors.split(','):
        gs.reset()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        res = surrogate.screened_da_scan(bt, gs, T, feature_turns=args.da_feature_turns,
                                         label_fraction=args.da_label_fraction,
                                         band_factor=3.0, band_sigma=args.da_band_sigma,
                                         epochs=args.da_epochs, seed=0, predictor=pred,
                                         normalisation=norm,
                                         progress=lambda what, d, n_: log(
                                             f'da: {pred} {what}, turn {d}, {n_} tracked'))
        wall = time.perf_counter() - t1
        ex = res['exact']
        assert np.array_equal(res['survival'][ex], plain['survival'][ex])
        errs = surrogate.da_errors(res['da'], plain['da'])
        out_scans[pred] = {
            'wall_s': round(wall, 2), 'speedup_vs_plain': plain['time'] / wall,
            'tracked_full_fraction': round(res['tracked_full_fraction'], 4),
            'band_particles': int(len(res['band'])),
            'band_stopped_above_a_loss': res['band_stopped_above_a_loss'],
            'parts_s': {k: round(v, 2) for k, v in res['time'].items()},
            'da_error_sigma_per_seed': [{k: round(v, 3) for k, v in e.items()} for e in errs],
            'da_error_sigma_max': max(e['max_abs'] for e in errs),
            'da_error_sigma_seed_mean_max': max(abs(e['mean_diff']) for e in errs),
            'da_error_sigma_seed_min_max': max(abs(e['min_diff']) for e in errs)}
        log(f'da: screened ({pred}) {wall:.1f} s, tracked {res["tracked_full_fraction"]:.3f}, '
            f'DA error: per angle max {out_scans[pred]["da_error_sigma_max"]:.2f}, '
            f'per seed (mean over angles) max '
            f'{out_scans[pred]["da_error_sigma_seed_mean_max"]:.3f}, per seed (minimum over '
            f'angles) max {out_scans[pred]["da_error_sigma_seed_min_max"]:.3f} sigma')
    da_plain = [float(np.mean(list(d.values()))) for d in plain['da']]
    da_plain_min = [float(min(d.values())) for d in plain['da']]
    out = {
        'metric': 'DA-scan wall time, plain vs surrogate-screened, 1 MI355X',
        'value': out_scans.get('mlp', next(iter(out_scans.values())))['speedup_vs_plain'],
        'unit': 'speed-up (plain time / screened time)', 'n_gpus': 1,
        'higher_is_better': True, 'vs_baseline': None,
        'dtype': 'f64 tracking, bf16/f32 network',
        'data': f'synthetic (polar grid, collider ring with seeded b3..b6 errors of '
                f'relative strength {args.da_error_scale:g})',
        'config': {'workload': f'DA scan of {len(lines)} error seeds of the collider ring, '
                               f'{args.da_angles} angles x '
                               f'{args.da_amplitudes if not args.da_grid_sizes else "uneven"} '
                               f'amplitudes (0-20 sigma) x 2 deltas per seed, {T} turns',
                   'particles': n, 'particles_per_seed': [int(v) for v in np.diff(gs.starts)],
                   'plain_s': round(plain['time'], 2),
                   'plain_mean_da_sigma_per_seed': [round(v, 3) for v in da_plain],
                   'plain_min_da_sigma_per_seed': [round(v, 3) for v in da_plain_min],
                   'screened': out_scans,
                   'feature_turns': args.da_feature_turns,
                   'label_fraction': args.da_label_fraction,
                   'band': {'factor': 3.0, 'sigma': args.da_band_sigma}},
    }
    print(json.dumps(out), flush=True)


def timed_steps(line, parts, steps, warmup, tps):
    """`warmup` untimed then `steps` timed launches of `tps` turns each on
    the current stream: (completed particle-turns, wall seconds, per-launch
    kernel ms from HIP events on the launch stream)."""
    import torch
    for _ in range(warmup):
        line.track(parts, num_turns=tps)
    torch.cuda.synchronize()
    at_turn0 = parts.at_turn.sum().item()
    stream = torch.cuda.current_stream()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        ev[k][0].record(stream)
        line.track(parts, num_turns=tps)
        ev[k][1].record(stream)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    pturns = parts.at_turn.sum().item() - at_turn0
    return pturns, elapsed, [a.elapsed_time(b) for a, b in ev]


def main():
    args = parse()
    if args.gpus != 1:
        print(f'bench.py: this build measures one GPU only (asked for {args.gpus}); '
              'particles shard across GPUs with no exchange during tracking, '
              'so N GPUs give N times the per-GPU rate.', file=sys.stderr)
        sys.exit(2)
    if args.workload == 'batch':
        main_batch(args)
        return
    if args.workload == 'da':
        main_da(args)
        return
    if args.workload == 'mlp':
        main_mlp(args)
        return
    traffic = None
    if args.traffic == 'measure':
        try:
            traffic = measure_traffic(args)
        except Exception as exc:      # counters unavailable: report null traffic
            log(f'traffic not measured ({exc})')

    import torch
    from xtrack_amd import lattices, optics, _native

    k2f, k2d = K2F, K2D
    if args.rematch:
        (k2f, k2d), _ = lattices.match_chromaticity(
            lambda a, b: lattices.collider_ring(k2f=a, k2d=b))

    def make_line():
        if args.ring_error_seed is not None:
            return lattices.collider_with_errors(args.ring_error_seed, k2f=k2f, k2d=k2d,
                                                 rel=args.ring_error_scale)
        return lattices.collider_ring(k2f=k2f, k2d=k2d)

    line = make_line()
    tracker = line.build_tracker(device='cuda:0',
                                 compile={'auto': 'auto', 'on': True, 'off': False}[args.compile],
                                 waves_per_eu=args.waves_per_eu, arithmetic=args.arithmetic)
    compiled = tracker.backend.is_compiled(tracker._handle)
    ref = line.particle_ref
    tw = optics.twiss(line, ref)
    chroma = optics.chromaticity(line, ref)
    parts = optics.matched_gaussian(line, ref, args.particles, NEMITT, NEMITT, SIGMA_Z,
                                    SIGMA_DELTA, seed=1, tw=tw)
    init = parts.copy()
    parts.move(_context='cuda:0')
    log(f'beam of {args.particles} particles on the GPU, line compiled: {compiled}')

    # algorithmic operations per particle-turn (kernel element code, counted)
    lib = _native.load()
    ops = np.zeros(6)
    coords = np.array([1e-5, 1e-7, 1e-5, 1e-7, 1e-3, 1e-5])
    _native.check(lib.xtm_line_count_ops(tracker._handle.ptr, coords.ctypes.data,
                                         float(ref._np('p0c')[0]), ref.mass0, ref.q0,
                                         ops.ctypes.data))
    flop_pt = ops[5] + FLOP_PER_CALL * ops[4]

    if args.copy_calibration:
        src = torch.ones(COPY_MIB * 2 ** 20 // 8, dtype=torch.float64, device='cuda:0')
        dst = torch.empty_like(src)
        torch.neg(src, out=dst)
        torch.cuda.synchronize()
        del src, dst

    tps = args.turns_per_step
    pturns, elapsed, kern_ms = timed_steps(line, parts, args.steps, args.warmup, tps)
    n_alive = int((parts.state > 0).sum().item())
    log(f'timed run ({args.arithmetic}): {np.mean(kern_ms):.2f} ms per step')

    value = pturns / elapsed
    # per launch: all particles alive at the start of the step x tps turns
    kern_s = float(np.mean(kern_ms)) / 1e3
    achieved = flop_pt * (pturns / args.steps) / kern_s / 1e12
    alg_bytes = STATE_BYTES_PER_PARTICLE * args.particles
    traffic_bytes = None
    if traffic is not None:
        traffic_bytes = traffic['FETCH_SIZE']['bytes'] + traffic['WRITE_SIZE']['bytes']
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
                               f'{len(line.elements)} elements, 7 TeV protons' +
                               ('' if args.ring_error_seed is None else
                                f', b3..b6 errors of seed {args.ring_error_seed} '
                                f'(relative {args.ring_error_scale:g})'),
                   'particles': args.particles, 'turns_per_step': tps,
                   'kernel': 'compiled line code' if compiled else 'element-program interpreter',
                   'particles_alive_end': n_alive,
                   'tunes': [round(float(tw['qx']), 4), round(float(tw['qy']), 4)],
                   'chromaticity': [round(float(chroma[0]), 3), round(float(chroma[1]), 3)]},
        'roofline': {'bound': 'fp64-valu', 'compute_pipe': 'FP64 VALU (MI355X FP64 peak '
                                                           '78.6 TFLOP/s; no matrix '
                                                           'instructions issued)',
                     'achieved': achieved, 'peak': PEAK_FP64_TFLOPS,
                     'unit': 'TFLOP/s', 'frac': achieved / PEAK_FP64_TFLOPS,
                     'traffic': traffic_bytes,
                     'algorithmic_bytes_per_launch': alg_bytes,
                     'traffic_counters': traffic,
                     'flop_per_particle_turn': flop_pt,
                     'ops_per_particle_turn': {'add': ops[0], 'mul': ops[1], 'div': ops[2],
                                               'sqrt': ops[3], 'calls': ops[4]},
                     'kernel_ms_mean': float(np.mean(kern_ms)),
                     'kernel_ms': [round(float(x), 3) for x in kern_ms]},
    }
    out['config']['arithmetic'] = args.arithmetic
    if not args.no_variants:
        other = 'fast' if args.arithmetic == 'exact' else 'exact'
        vline = make_line()
        vline.build_tracker(device='cuda:0', waves_per_eu=args.waves_per_eu, arithmetic=other)
        vparts = init.copy(_context='cuda:0')
        vpt, vel, vms = timed_steps(vline, vparts, args.steps, args.warmup, tps)
        vach = flop_pt * (vpt / args.steps) / (float(np.mean(vms)) / 1e3) / 1e12
        out['variants'] = {other: {
            'value': vpt / vel, 'ms_per_step': vel / args.steps * 1e3,
            'kernel_ms_mean': float(np.mean(vms)),
            'roofline_frac': vach / PEAK_FP64_TFLOPS, 'achieved': vach,
            'same_beam_and_steps': True}}
        log(f'timed run ({other}): {np.mean(vms):.2f} ms per step')
        del vparts
    if not args.no_batch:
        out['batch'] = batch_leg(args, steps=2, kernels=('compiled', 'interpreter'))
    if not args.no_cpu_baseline:
        out['parity'], out['cpu_baseline'] = parity_and_cpu_legs(
            args, make_line, parts, init, (args.warmup + args.steps) * tps)
    log('done')
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
