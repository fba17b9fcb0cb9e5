"""Batched lattice variants on the GPU (xtm_batch_track through
BatchTracker): every variant's particles must end bit for bit as the CPU
oracle leaves them when it tracks that variant alone, for uneven particle
counts (an empty variant included), launches split into turn blocks, HIP
graph replays and every kernel instance the autotuner chooses from."""
import numpy as np
import pytest

import xtrack_amd as xt
from xtrack_amd import lattices
from helpers import mixed_line, random_particles, compare_exact
from oracle_backend import OracleBackend

pytestmark = pytest.mark.gpu

COUNTS = (700, 37, 0, 1300, 256, 5)


def _variants():
    out = [lattices.collider_variant(s, n_arcs=8, cells_per_arc=(1, 3), rel_err=1e-3)
           for s in range(4)]
    out.insert(2, mixed_line(n_cells=2))
    out.append(mixed_line(n_cells=1, aperture=0.01))
    return out[:len(COUNTS)]


def _collider_beam(n, seed):
    """Small transverse and momentum spread, with one particle in ten at
    5-25 mm so that some are lost at the apertures."""
    rng = np.random.default_rng(seed)
    x = rng.normal(0, 5e-4, n)
    far = rng.random(n) < 0.1
    x[far] = rng.uniform(5e-3, 2.5e-2, far.sum()) * rng.choice([-1, 1], far.sum())
    return xt.Particles(p0c=7e12, x=x, px=rng.normal(0, 5e-6, n), y=rng.normal(0, 5e-4, n),
                        py=rng.normal(0, 5e-6, n), zeta=rng.normal(0, 0.05, n),
                        delta=rng.normal(0, 1e-4, n))


def _particles():
    out = []
    for k, n in enumerate(COUNTS):
        if n == 0:
            out.append(xt.Particles(p0c=7e12, x=np.zeros(0)))
        elif k in (2, 5):
            out.append(random_particles(n, seed=100 + k, scale=4e-3, p0c=7e9))
        else:
            out.append(_collider_beam(n, seed=100 + k))
    return out


def _oracle(lines, parts, num_turns):
    res = []
    for ln, pp in zip(lines, parts):
        q = pp.copy()
        if q._capacity:
            ln.build_tracker(backend=OracleBackend())
            ln.track(q, num_turns=num_turns)
        res.append(q)
    return res


def _same_species(parts):
    """The batch needs one reference species and t_sim for all sets; the
    sets above differ in p0c only, a per-particle field."""
    for pp in parts:
        pp.q0, pp.mass0, pp.t_sim = 1.0, xt.PROTON_MASS_EV, 0.0
    return parts


@pytest.mark.parametrize('compiled', [True, False], ids=['compiled', 'interpreter'])
def test_batch_bit_exact_uneven_variants(compiled):
    print(f'[batch] uneven variants, turn blocks, compiled={compiled}', flush=True)
    lines = _variants()
    parts = _same_species(_particles())
    ref = _oracle(_variants(), [p.copy() for p in parts], 5)
    bt = xt.BatchTracker(lines, device='cuda:0', compile=compiled)
    assert bt.is_compiled == compiled
    got = [p.copy() for p in parts]
    info = bt.track(got, num_turns=5, turns_per_launch=2)
    assert info['launches'] == 3
    lost = 0
    for k, (g, r) in enumerate(zip(got, ref)):
        if COUNTS[k] == 0:
            assert g._capacity == 0
            continue
        bad = compare_exact(g, r)
        assert not bad, (k, bad)
        lost += int((r.sorted_by_id()['state'] <= 0).sum())
    assert lost > 0
    print(f'lost={lost}', flush=True)


@pytest.mark.parametrize('block', [64, 128])
def test_batch_block_sizes_bit_exact(block):
    """The compiled batch code and the interpreter after it at 64 and 128
    threads per block (the per-thread LDS slots sized by the launch):
    every variant as the oracle leaves it."""
    print(f'[batch] block size {block}', flush=True)
    lines = _variants()
    parts = _same_species(_particles())
    ref = _oracle(_variants(), [p.copy() for p in parts], 4)
    bt = xt.BatchTracker(lines, device='cuda:0', compile=True)
    bt.config = (block, 5)
    got = [p.copy() for p in parts]
    bt.track(got, num_turns=4, turns_per_launch=2)
    assert int(bt.blocks()[:, 2].max()) <= block
    for k, (g, r) in enumerate(zip(got, ref)):
        if COUNTS[k]:
            bad = compare_exact(g, r)
            assert not bad, (k, bad)


@pytest.mark.parametrize('compiled', [False, True], ids=['interpreter', 'compiled'])
def test_batch_graph_replays_match_eager(compiled):
    """One launch of 2 turns captured in a HIP graph and replayed three
    times, with eager work and synchronisation between the replays; after
    each replay the particles equal those of the same number of eager
    launches."""
    import torch
    print(f'[batch] graph replays, compiled={compiled}', flush=True)
    lines = _variants()
    parts = _same_species(_particles())
    bt = xt.BatchTracker(lines, device='cuda:0', compile=compiled)
    assert bt.is_compiled == compiled
    big, starts = bt._concat(parts)
    bt._set_layout(starts, (256, 4))
    eager = big.copy()
    graph_p = big.copy()
    warm = big.copy()
    bt._launch(warm, 1)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        bt._launch(graph_p, 2)
    torch.cuda.synchronize()
    for r in range(3):
        bt._launch(eager, 2)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        bad = compare_exact(graph_p, eager)
        assert not bad, (r, bad)
    # and the graph path of track() gives the eager result
    a = [p.copy() for p in parts]
    b = [p.copy() for p in parts]
    bt.track(a, num_turns=7, turns_per_launch=3)
    info = bt.track(b, num_turns=7, turns_per_launch=3, use_graph=True)
    assert info['graph']
    for k in range(len(parts)):
        if COUNTS[k]:
            assert not compare_exact(a[k], b[k]), k


def test_batch_tuner_kernel_instances_agree():
    """Every kernel instance gives the same bits; the tuner keeps one of
    them."""
    print('[batch] tuner', flush=True)
    lines = _variants()
    parts = _same_species(_particles())
    bt = xt.BatchTracker(lines, device='cuda:0', compile=False)
    cfg = bt.tune(parts, turns=1)
    assert cfg in xt.batch.CANDIDATES and set(bt.tune_timings) == set(xt.batch.CANDIDATES)
    assert {c[0] for c in xt.batch.CANDIDATES} == {64, 128, 256}
    results = {}
    for cand in xt.batch.CANDIDATES:
        bt.config = cand
        got = [p.copy() for p in parts]
        bt.track(got, num_turns=3)
        results[cand] = got
    base = results[xt.batch.CANDIDATES[0]]
    for cand, got in results.items():
        for k in range(len(parts)):
            if COUNTS[k]:
                assert not compare_exact(got[k], base[k]), (cand, k)


def test_batch_rejects_mixed_species():
    lines = _variants()[:2]
    bt = xt.BatchTracker(lines, device='cuda:0')
    a = random_particles(10, seed=1)
    b = random_particles(10, seed=2)
    b.q0 = 2.0
    with pytest.raises(ValueError):
        bt.track([a, b], num_turns=1)


def test_batch_full_shape_sample_bit_exact():
    """The batched workload at its full shape: 64 variants of the collider
    ring (seeds 0..63, 8-40 cells per arc, quadrupole and sextupole
    strengths scattered by N(0, 1e-4)), particle counts log-uniform in
    1,000-200,000 (about 2.4 million particles), two turns in one launch
    on the compiled batch code.  Per variant, 16 particles plus every lost
    one (at most 64) are tracked alone on the oracle: identical bits."""
    from xtrack_amd import lattices
    print('[batch] 64 variants, full shape', flush=True)
    rng = np.random.default_rng(12345)
    counts = np.exp(rng.uniform(np.log(1000), np.log(200_000), 64)).astype(np.int64)
    lines = [lattices.collider_variant(s, k2f=0.0319663, k2d=-0.06119144) for s in range(64)]
    parts = []
    for k, n in enumerate(counts):
        r = np.random.default_rng(1000 + k)
        x = r.normal(0, 3e-4, n)
        far = r.random(n) < 0.002
        x[far] = r.uniform(0.012, 0.03, far.sum()) * r.choice([-1, 1], far.sum())
        parts.append(xt.Particles(p0c=7e12, mass0=xt.PROTON_MASS_EV, x=x,
                                  px=r.normal(0, 3e-6, n), y=r.normal(0, 3e-4, n),
                                  py=r.normal(0, 3e-6, n), zeta=r.normal(0, 0.0755, n),
                                  delta=r.normal(0, 1.13e-4, n)))
    init = [p.copy() for p in parts]
    bt = xt.BatchTracker(lines, device='cuda:0', compile=True)
    assert bt.is_compiled
    print(f'[batch] {int(counts.sum())} particles, tracking', flush=True)
    bt.track(parts, num_turns=2, turns_per_launch=2)
    n_lost_all = 0
    for k in range(64):
        g = parts[k].to_numpy()
        lost = np.flatnonzero(g['state'] <= 0)
        n_lost_all += len(lost)
        sample = np.union1d(np.random.default_rng(k).choice(counts[k], 16, replace=False),
                            lost[:64])
        h = init[k].to_numpy()
        ps = xt.Particles(p0c=7e12, mass0=xt.PROTON_MASS_EV,
                          **{ff: h[ff][sample] for ff in ('x', 'px', 'y', 'py', 'zeta',
                                                           'delta')},
                          particle_id=sample)
        ln = lattices.collider_variant(k, k2f=0.0319663, k2d=-0.06119144)
        ln.build_tracker(backend=OracleBackend(num_threads=8))
        ln.track(ps, num_turns=2)
        got = xt.Particles(p0c=7e12, mass0=xt.PROTON_MASS_EV, x=np.zeros(len(sample)))
        got._set_all({nn: np.ascontiguousarray(vv[sample]) for nn, vv in g.items()})
        bad = compare_exact(got, ps)
        assert not bad, (k, bad)
    print(f'[batch] lost {n_lost_all}', flush=True)
    assert n_lost_all > 1000
