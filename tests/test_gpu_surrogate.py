"""The surrogate network's matrix-core products and gradients (xtm_gemm_bf16
through xtrack_amd.surrogate) against CPU references in float64 with the
same BF16 rounding points: plain and transposed operands, ragged sizes,
and long-k products that split k over blocks (weight gradients)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bf16_round(a):
    """float64 values of float32 `a` rounded to BF16 (nearest even)."""
    u = np.asarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


@pytest.mark.parametrize('M,N,K,ta,tb', [(1, 1, 1, False, False), (70, 33, 17, False, False),
                                         (130, 257, 16, True, False),
                                         (64, 64, 300, False, True), (5, 100, 40, True, True),
                                         (256, 256, 4096, True, False),
                                         (16, 256, 1500, True, False),
                                         (2000, 256, 256, False, False),
                                         (300, 1, 256, False, False)])
def test_gemm_against_cpu(M, N, K, ta, tb):
    import torch
    from xtrack_amd import surrogate as S
    print(f'[gemm] {M}x{N}x{K} ta={ta} tb={tb}', flush=True)
    rng = np.random.default_rng(M * 1000 + N)
    a = rng.normal(size=(K, M) if ta else (M, K)).astype(np.float32)
    b = rng.normal(size=(N, K) if tb else (K, N)).astype(np.float32)
    bias = rng.normal(size=N).astype(np.float32)
    dev = torch.device('cuda:0')
    ab = S.to_bf16(torch.from_numpy(a).to(dev))
    bb = S.to_bf16(torch.from_numpy(b).to(dev))
    for relu, use_bias in ((False, False), (True, True)):
        c, cb = S.gemm(ab, bb, ta, tb, bias=torch.from_numpy(bias).to(dev) if use_bias else None,
                       relu=relu, want_bf16=True)
        A, B = _bf16_round(a), _bf16_round(b)
        ref = (A.T if ta else A) @ (B.T if tb else B)
        if use_bias:
            ref = ref + bias
        if relu:
            ref = np.maximum(ref, 0.0)
        got = c.cpu().numpy().astype(np.float64)
        scale = np.abs(A).max() * np.abs(B).max() * K + 1.0
        assert np.max(np.abs(got - ref)) <= 1e-6 * scale, np.max(np.abs(got - ref))
        assert np.array_equal(cb.cpu().numpy(), S.to_bf16(c).cpu().numpy())


def test_mlp_gradients_against_cpu():
    """Backward pass of the network (matrix-core products) against float64
    backpropagation on the CPU with the same BF16 roundings of activations,
    weights and output gradients."""
    import torch
    from xtrack_amd import surrogate as S
    print('[mlp gradients]', flush=True)
    net = S.MLP(sizes=(16, 64, 48, 32, 1), seed=3)
    rng = np.random.default_rng(1)
    x = torch.from_numpy(rng.normal(size=(200, 16)).astype(np.float32)).to('cuda:0')
    y = torch.from_numpy(rng.normal(size=200).astype(np.float32)).to('cuda:0')
    out = net.forward(x, keep=True)
    dout = 2 * (out - y) / 200
    gW, gb = net.backward(dout)
    # CPU reference
    W = [w.cpu().numpy() for w in net.W]
    b = [v.cpu().numpy().astype(np.float64) for v in net.b]
    h = [_bf16_round(x.cpu().numpy())]
    zs = []
    for li in range(4):
        z = (h[-1] @ _bf16_round(W[li]) + b[li]).astype(np.float32)
        if li < 3:
            z = np.maximum(z, 0)
            h.append(_bf16_round(z))
        zs.append(z.astype(np.float64))
    assert np.allclose(out.cpu().numpy(), zs[-1][:, 0], rtol=1e-5, atol=1e-5)
    dz = (2 * (zs[-1][:, 0].astype(np.float32) - y.cpu().numpy()) / 200).reshape(-1, 1)
    for li in range(3, -1, -1):
        dzb = _bf16_round(dz.astype(np.float32))
        gw_ref = h[li].T @ dzb
        assert np.allclose(gW[li].cpu().numpy(), gw_ref, rtol=2e-3, atol=1e-5 * np.abs(gw_ref).max()), li
        assert np.allclose(gb[li].cpu().numpy(), dz.sum(axis=0), rtol=1e-4, atol=1e-6), li
        if li > 0:
            dh = dzb @ _bf16_round(W[li]).T
            dz = (dh * (zs[li - 1] > 0)).astype(np.float32)


def test_mlp_fits_a_smooth_function():
    """Training on the matrix cores lowers the loss of a smooth target by
    orders of magnitude."""
    import torch
    from xtrack_amd import surrogate as S
    print('[mlp fit]', flush=True)
    rng = np.random.default_rng(5)
    x = rng.uniform(-1, 1, size=(4000, 16)).astype(np.float32)
    y = (np.sin(2 * x[:, 0]) + x[:, 1] * x[:, 2] + 0.5 * x[:, 3] ** 2).astype(np.float32)
    dev = torch.device('cuda:0')
    net = S.MLP(seed=1)
    hist = net.fit(torch.from_numpy(x).to(dev), torch.from_numpy(y).to(dev), epochs=60,
                   batch=256, lr=2e-3)
    assert hist[-1] < 0.05 * hist[0], hist[::10]
    pred = net.predict(torch.from_numpy(x[:500]).to(dev)).cpu().numpy()
    assert np.sqrt(np.mean((pred - y[:500]) ** 2)) < 0.15 * np.std(y)


def _sext_ring():
    from xtrack_amd import lattices
    return lattices.sextupole_fodo_ring()


def test_surrogate_da_scan_against_plain():
    """The surrogate-screened scan on a sextupole-perturbed ring: every
    particle it tracks in full has the survival turn of plain tracking (the
    feature turns and the rest are one continued track), only part of the
    grid is tracked in full, and the per-angle apertures stay within one
    grid step of plain tracking's."""
    import time
    import torch
    from xtrack_amd import da, surrogate as S
    print('[surrogate da scan]', flush=True)
    kw = dict(r_max=12, n_r=48, n_angles=9, num_turns=2000, nemitt_x=2.5e-6,
              nemitt_y=2.5e-6, deltas=(0.0, 2e-3))
    line = _sext_ring()
    line.build_tracker(device='cuda:0')
    t0 = time.perf_counter()
    plain = da.da_scan(line, **kw)
    t_plain = time.perf_counter() - t0
    res = S.surrogate_da_scan(line, feature_turns=256, train_fraction=0.2, band=1.5,
                              epochs=120, seed=0, **kw)
    ex = res['exact']
    assert np.array_equal(res['survival'][ex], plain['survival'][ex])
    frac = len(res['tracked_full']) / len(res['r'])
    step = kw['r_max'] / kw['n_r']
    errs = [abs(res['da'][k] - plain['da'][k]) for k in plain['da']]
    print(f'tracked in full {frac:.2f}, DA errors max {max(errs):.3f} mean '
          f'{np.mean(errs):.3f} sigma, plain {t_plain:.2f} s, surrogate '
          f'{res["time"]}', flush=True)
    assert frac < 0.7
    assert max(errs) <= step + 1e-9
    assert res['loss_history'][-1] < res['loss_history'][0]


def test_network_gain_over_interpolation():
    """What the trained network gains: the same screened scan (feature
    turns, 5 % of the survivors labelled, boundary band tracked in full)
    with the network replaced by interpolating the labels in amplitude
    within each angle (predictor='interp').  Both scans' tracked particles
    keep their plain survival turns; the network's apertures are within two
    grid steps of plain tracking and no worse than the interpolation's."""
    import torch
    import xtrack_amd as xt
    from xtrack_amd import surrogate as S
    print('[network gain]', flush=True)
    kw = dict(r_max=12, n_r=60, n_angles=12, deltas=(0.0, 2e-3), nemitt_x=2.5e-6,
              nemitt_y=2.5e-6)
    line = _sext_ring()
    line.build_tracker(device='cuda:0')
    dev = torch.device('cuda:0')
    bt = xt.BatchTracker([line], device=dev, compile=True)
    gs = S.GridSet([line], device=dev, **kw)
    T = 2000
    plain = S.plain_da_scan_batch(bt, gs, T)
    out = {}
    for pred in ('mlp', 'interp'):
        gs.reset()
        res = S.screened_da_scan(bt, gs, T, feature_turns=128, label_fraction=0.05,
                                 band_factor=2.0, band_sigma=0.5, epochs=100, seed=0,
                                 predictor=pred)
        ex = res['exact']
        assert np.array_equal(res['survival'][ex], plain['survival'][ex])
        err = S.da_errors(res['da'], plain['da'])[0]
        out[pred] = (err['max_abs'], res['tracked_full_fraction'])
        print(f'{pred}: DA error max {err["max_abs"]:.3f} sigma (mean diff '
              f'{err["mean_diff"]:+.3f}), tracked in full {res["tracked_full_fraction"]:.3f}',
              flush=True)
    step = kw['r_max'] / kw['n_r']
    assert out['mlp'][0] <= 2 * step + 1e-9
    assert out['mlp'][0] <= out['interp'][0] + 1e-9


def test_batched_screened_scan_uneven_error_seeds():
    """The screened scan under batching: three b3..b6 error seeds of the
    collider ring (30 units) with uneven amplitude grids, all in one batch
    tracker.  Every particle the scan tracks in full keeps exactly its
    plain-scan survival turn, each seed's grid has its own size, and each
    seed's aperture (minimum over angles and deltas) is within two grid
    steps of plain tracking."""
    import torch
    import xtrack_amd as xt
    from xtrack_amd import lattices, surrogate as S
    print('[batched screened scan]', flush=True)
    lines = [lattices.collider_with_errors(s, k2f=0.0319663, k2d=-0.06119144, rel=3e-3)
             for s in range(3)]
    dev = torch.device('cuda:0')
    bt = xt.BatchTracker(lines, device=dev, compile=True)
    n_r = [40, 90, 60]
    gs = S.GridSet(lines, 20.0, n_r, 8, (0.0, 2.7e-4), 3.75e-6, 3.75e-6, dev)
    assert list(np.diff(gs.starts)) == [n * 8 * 2 for n in n_r]
    T = 600
    plain = S.plain_da_scan_batch(bt, gs, T)
    assert (plain['survival'] < T).any() and (plain['survival'] == T).any()
    gs.reset()
    res = S.screened_da_scan(bt, gs, T, feature_turns=64, label_fraction=0.05,
                             band_factor=3.0, band_sigma=0.5, epochs=60, seed=0,
                             normalisation=S.normal_form_4d(lines[0], dev))
    ex = res['exact']
    assert np.array_equal(res['survival'][ex], plain['survival'][ex])
    errs = S.da_errors(res['da'], plain['da'])
    print('tracked in full %.3f, min-over-angles DA errors %s' % (
        res['tracked_full_fraction'], [round(e['min_diff'], 3) for e in errs]), flush=True)
    for l, e in enumerate(errs):
        assert abs(e['min_diff']) <= 2 * 20.0 / n_r[l] + 1e-9, (l, e)


def test_fit_graph_replays_equal_eager_steps():
    """The training step captured once in a HIP graph and replayed per
    minibatch (with the minibatch copies, a short last minibatch run eagerly
    and the per-epoch loss read back between replays) gives the eager fit's
    weights, biases and losses bit for bit, censored targets included."""
    import torch
    from xtrack_amd import surrogate as S
    print('[fit graph]', flush=True)
    dev = torch.device('cuda:0')
    g = torch.Generator(device='cpu').manual_seed(3)
    x = torch.randn(3000, 16, generator=g).to(dev)
    y = (x[:, :4].sum(dim=1) + 0.1 * torch.randn(3000, generator=g).to(dev)).contiguous()
    cens = (torch.rand(3000, generator=g) < 0.3).to(dev)
    nets, hists = [], []
    for graph in (False, True):
        net = S.MLP(device=dev, seed=0)
        hists.append(net.fit(x, y, epochs=4, batch=1024, seed=0, censored=cens, graph=graph))
        nets.append(net)
    assert hists[0] == hists[1], hists
    for a, b in zip(nets[0].W + nets[0].b, nets[1].W + nets[1].b):
        assert torch.equal(a, b)
    assert hists[0][-1] < hists[0][0]
