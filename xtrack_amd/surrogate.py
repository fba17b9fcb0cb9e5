"""Survival-turn surrogate for dynamic-aperture scans (no counterpart in
the reference).

A small MLP (16 -> 256 -> 256 -> 256 -> 1, ReLU) predicts log(survival
turn) of a particle from its initial normalised coordinates and features of
its first turns (tracked by the plain tracker).  Its products run on the
matrix cores through xtm_gemm_bf16 (BF16 operands, FP32 accumulation),
forward and backward; the FP32 master weights, the bias and loss sums and
the Adam update are torch tensor operations on the same device.

The scan (``surrogate_da_scan``) tracks every grid particle for the feature
turns, trains the network on a tracked subset, and tracks to the full turn
count only the particles whose predicted survival lies within a factor of
the target (the boundary band); particles the network places well inside
or outside are assigned their prediction.  Plain full tracking of the same
grid is the reference the scan's aperture error is measured against.
"""
import ctypes
import math
import time

import numpy as np

from . import _native

SIZES = (16, 256, 256, 256, 1)


def _t():
    import torch
    return torch


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream(dev):
    torch = _t()
    return ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)


def to_bf16(x):
    """BF16 copy of an FP32 tensor (round to nearest even), as int16 bits."""
    torch = _t()
    return x.to(torch.bfloat16).view(torch.int16)


def gemm(a, b, trans_a=False, trans_b=False, bias=None, relu=False, want_bf16=False):
    """op(a) @ op(b) on the matrix cores: a, b are BF16 bit tensors (int16,
    2-D, contiguous); returns (FP32 result, BF16 bits of it or None)."""
    torch = _t()
    lib = _native.load()
    if a.dtype != torch.int16 or b.dtype != torch.int16 or a.dim() != 2 or b.dim() != 2:
        raise ValueError('gemm operands are 2-D BF16 bit tensors')
    a, b = a.contiguous(), b.contiguous()
    M, K = (a.shape[1], a.shape[0]) if trans_a else (a.shape[0], a.shape[1])
    Kb, N = (b.shape[1], b.shape[0]) if trans_b else (b.shape[0], b.shape[1])
    if K != Kb:
        raise ValueError(f'inner dimensions differ: {K} and {Kb}')
    sam, sak = (1, a.shape[1]) if trans_a else (a.shape[1], 1)
    sbk, sbn = (1, b.shape[1]) if trans_b else (b.shape[1], 1)
    if bias is not None and (bias.dtype != torch.float32 or bias.numel() != N):
        raise ValueError('bias must be FP32 with one value per output column')
    c = torch.empty((M, N), dtype=torch.float32, device=a.device)
    cb = torch.empty((M, N), dtype=torch.int16, device=a.device) if want_bf16 else None
    _native.check(lib.xtm_gemm_bf16(_ptr(a), sam, sak, _ptr(b), sbk, sbn, _ptr(c), N,
                                    _ptr(cb), _ptr(bias.contiguous() if bias is not None
                                                   else None),
                                    int(bool(relu)), M, N, K, _stream(a.device)))
    return c, cb


class MLP:
    """16 -> 256 -> 256 -> 256 -> 1 with ReLU; FP32 master weights, BF16
    products on the matrix cores."""

    def __init__(self, device='cuda:0', sizes=SIZES, seed=0):
        torch = _t()
        self.device = torch.device(device)
        self.sizes = tuple(sizes)
        g = torch.Generator(device='cpu').manual_seed(seed)
        self.W, self.b = [], []
        for i, o in zip(self.sizes[:-1], self.sizes[1:]):
            w = torch.randn((i, o), generator=g, dtype=torch.float32) * math.sqrt(2.0 / i)
            self.W.append(w.to(self.device))
            self.b.append(torch.zeros(o, dtype=torch.float32, device=self.device))
        self._adam = None
        self.x_mean = self.x_std = None
        self.y_mean, self.y_std = 0.0, 1.0

    def forward(self, x, keep=False):
        """x: (B, 16) FP32 on the device (already normalised)."""
        h = to_bf16(x)
        acts = [h]
        zs = []
        n = len(self.W)
        for li in range(n):
            last = li == n - 1
            z, zb = gemm(h, to_bf16(self.W[li]), bias=self.b[li], relu=not last,
                         want_bf16=not last)
            zs.append(z)
            h = zb
            if not last:
                acts.append(h)
        if keep:
            self._cache = (acts, zs)
        return zs[-1][:, 0]

    def backward(self, dout):
        """Gradients of the weights and biases for d(loss)/d(output) =
        dout (B,), after forward(keep=True)."""
        acts, zs = self._cache
        n = len(self.W)
        dz = dout.reshape(-1, 1).to(_t().float32)
        gW, gb = [None] * n, [None] * n
        for li in range(n - 1, -1, -1):
            dzb = to_bf16(dz)
            gW[li], _ = gemm(acts[li], dzb, trans_a=True)       # h^T dz
            gb[li] = dz.sum(dim=0)
            if li > 0:
                dh, _ = gemm(dzb, to_bf16(self.W[li]), trans_b=True)   # dz W^T
                dz = dh * (zs[li - 1] > 0).to(dh.dtype)
        return gW, gb

    def adam_step(self, gW, gb, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        torch = _t()
        if self._adam is None:
            self._adam = {'t': 0, 'm': [torch.zeros_like(p) for p in self.W + self.b],
                          'v': [torch.zeros_like(p) for p in self.W + self.b]}
        st = self._adam
        st['t'] += 1
        params, grads = self.W + self.b, gW + gb
        for k, (p, g) in enumerate(zip(params, grads)):
            st['m'][k].mul_(beta1).add_(g, alpha=1 - beta1)
            st['v'][k].mul_(beta2).addcmul_(g, g, value=1 - beta2)
            mh = st['m'][k] / (1 - beta1 ** st['t'])
            vh = st['v'][k] / (1 - beta2 ** st['t'])
            p.sub_(lr * mh / (vh.sqrt() + eps))

    def fit(self, x, y, epochs=200, batch=1024, lr=2e-3, seed=0):
        """Mean-squared error on standardised inputs and targets; x (N, 16),
        y (N,) torch FP32 on the device.  Returns the loss per epoch."""
        torch = _t()
        self.x_mean = x.mean(dim=0)
        self.x_std = x.std(dim=0).clamp_min(1e-12)
        self.y_mean, self.y_std = float(y.mean()), float(y.std().clamp_min(1e-12))
        xn = (x - self.x_mean) / self.x_std
        yn = (y - self.y_mean) / self.y_std
        g = torch.Generator(device='cpu').manual_seed(seed)
        hist = []
        n = len(xn)
        for _ in range(epochs):
            perm = torch.randperm(n, generator=g).to(self.device)
            tot = 0.0
            for s in range(0, n, batch):
                idx = perm[s:s + batch]
                out = self.forward(xn[idx], keep=True)
                err = out - yn[idx]
                gW, gb = self.backward(2.0 * err / len(idx))
                self.adam_step(gW, gb, lr=lr)
                tot += float((err * err).sum())
            hist.append(tot / n)
        return hist

    def predict(self, x, batch=65536):
        torch = _t()
        xn = (x - self.x_mean) / self.x_std
        out = [self.forward(xn[s:s + batch]) for s in range(0, len(xn), batch)]
        return torch.cat(out) * self.y_std + self.y_mean


# ---------------------------------------------------------------------------
# dynamic-aperture scan screened by the surrogate

def _take(p, idx):
    """Particles holding the slots `idx` (device index tensor) of `p`."""
    from .particles import Particles
    out = Particles.__new__(Particles)
    out._capacity = int(len(idx))
    out.q0, out.mass0, out.t_sim = p.q0, p.mass0, p.t_sim
    out.start_tracking_at_element = p.start_tracking_at_element
    torch = _t()

    def take(vv):
        if vv.dtype == torch.uint32:     # no unsigned indexing on the device
            return vv.view(torch.int32)[idx].contiguous().view(torch.uint32)
        return vv[idx].contiguous()

    out._arrays = {nn: take(vv) for nn, vv in p._arrays.items()}
    out._device = p._device
    return out


def track_features(line_or_batch, particles, num_turns, inv_w4, co, sx, sy, starts=None):
    """Track `particles` (a device particle set) for num_turns whole turns
    with the feature accumulators updated inside the tracking kernel at the
    start of every turn (xtm_track_line_features / xtm_batch_track_features);
    returns the accumulators, a (XTM_FEAT_N, capacity) float64 tensor on the
    device, row k = accumulator k of every particle slot.

    line_or_batch: a Line with a device tracker, or a BatchTracker together
    with `starts` (the variants' particle ranges in `particles`)."""
    torch = _t()
    lib = _native.load()
    n = particles._capacity
    acc = torch.zeros((_native.XTM_FEAT_N, n), dtype=torch.float64,
                      device=particles._arrays['x'].device)
    t0 = int(particles._arrays['at_turn'].min().item()) if n else 0
    fst = _native.make_features_struct(acc, n, t0, t0 + num_turns, np.asarray(inv_w4),
                                       co, (1.0 / sx, 1.0 / sy))
    pst = _native.make_particles_struct(particles)
    if starts is None:
        tr = line_or_batch.tracker
        _native.check(lib.xtm_track_line_features(
            tr._handle.ptr, ctypes.byref(pst), int(num_turns), ctypes.byref(fst),
            ctypes.c_uint64(int(tr.track_flags.get_flags_register())),
            tr.backend.stream()))
    else:
        bt = line_or_batch
        bt._set_layout(np.asarray(starts, dtype=np.int64), bt.config or (256, 4))
        _native.check(lib.xtm_batch_track_features(
            bt._batch, ctypes.byref(pst), int(num_turns), ctypes.byref(fst),
            ctypes.c_uint64(int(bt.track_flags.get_flags_register())), bt._stream()))
    return acc


def features_from_acc(acc, x_norm0, y_norm0, delta):
    """The 16 network inputs per particle from the in-kernel accumulators:
    r, theta, delta, the initial actions, mean, spread and maximum of the
    normalised actions over the recorded turns (relative to the initial
    ones), the mean phase advance per turn in each plane (tunes), the
    fraction of the turns recorded alive, cos and sin of theta."""
    torch = _t()
    dev = acc.device
    n = acc[0].clamp_min(1.0)       # turns recorded
    nturn = float(acc[0].max().clamp_min(1.0))
    mjx, mjy = acc[1] / n, acc[4] / n
    sjx = torch.sqrt((acc[2] / n - mjx * mjx).clamp_min(0.0))
    sjy = torch.sqrt((acc[5] / n - mjy * mjy).clamp_min(0.0))
    j0x, j0y = acc[13].clamp_min(1e-12), acc[14].clamp_min(1e-12)
    nph = (acc[0] - 1.0).clamp_min(1.0)
    qx, qy = acc[7] / nph / (2 * math.pi), acc[8] / nph / (2 * math.pi)
    r0 = torch.sqrt(torch.as_tensor(x_norm0 ** 2 + y_norm0 ** 2, device=dev))
    th0 = torch.atan2(torch.as_tensor(y_norm0, device=dev), torch.as_tensor(x_norm0, device=dev))
    f = torch.stack([r0, th0, torch.as_tensor(delta * 1e3, device=dev),
                     torch.log1p(j0x), torch.log1p(j0y),
                     torch.log(mjx.clamp_min(1e-12) / j0x), torch.log(mjy.clamp_min(1e-12) / j0y),
                     sjx / mjx.clamp_min(1e-12), sjy / mjy.clamp_min(1e-12),
                     torch.log(acc[3].clamp_min(1e-12) / j0x),
                     torch.log(acc[6].clamp_min(1e-12) / j0y),
                     qx, qy, acc[0] / nturn, torch.cos(th0), torch.sin(th0)], dim=1)
    return torch.nan_to_num(f, nan=0.0, posinf=50.0, neginf=-50.0).to(torch.float32)


def features(mon, x_norm0, y_norm0, delta, invW4, sx, sy, n_turns):
    """16 features per particle from its initial normalised position and
    the turn-by-turn record of its first n_turns turns: r, theta, delta,
    the initial actions, mean, spread and maximum of the normalised actions
    over the turns (relative to the initial ones), the mean phase advance
    per turn in each plane (tunes), the fraction of the turns survived."""
    torch = _t()
    X = torch.stack([mon.x, mon.px, mon.y, mon.py])            # 4 x N x T
    st = mon.state
    alive = (st > 0).to(torch.float64)
    Wi = torch.as_tensor(invW4, dtype=torch.float64, device=X.device)
    Z = torch.einsum('ij,jnt->int', Wi, X.to(torch.float64))
    zx, zpx = Z[0] / sx, Z[1] / sx
    zy, zpy = Z[2] / sy, Z[3] / sy
    jx, jy = 0.5 * (zx * zx + zpx * zpx), 0.5 * (zy * zy + zpy * zpy)
    nal = alive.sum(dim=1).clamp_min(1.0)
    r0 = torch.sqrt(torch.as_tensor(x_norm0 ** 2 + y_norm0 ** 2, device=X.device))
    th0 = torch.atan2(torch.as_tensor(y_norm0, device=X.device),
                      torch.as_tensor(x_norm0, device=X.device))
    j0x, j0y = jx[:, 0].clamp_min(1e-12), jy[:, 0].clamp_min(1e-12)
    mjx = (jx * alive).sum(1) / nal
    mjy = (jy * alive).sum(1) / nal
    sjx = torch.sqrt(((jx - mjx[:, None]) ** 2 * alive).sum(1) / nal)
    sjy = torch.sqrt(((jy - mjy[:, None]) ** 2 * alive).sum(1) / nal)
    mxx = (jx * alive).amax(1)
    mxy = (jy * alive).amax(1)

    def tune(z, pz):
        ph = torch.atan2(-pz, z)
        d = torch.remainder(ph[:, 1:] - ph[:, :-1], 2 * math.pi)
        a = alive[:, 1:]
        return (d * a).sum(1) / a.sum(1).clamp_min(1.0) / (2 * math.pi)

    f = torch.stack([r0, th0, torch.as_tensor(delta * 1e3, device=X.device),
                     torch.log1p(j0x), torch.log1p(j0y),
                     torch.log(mjx.clamp_min(1e-12) / j0x), torch.log(mjy.clamp_min(1e-12) / j0y),
                     sjx / mjx.clamp_min(1e-12), sjy / mjy.clamp_min(1e-12),
                     torch.log(mxx.clamp_min(1e-12) / j0x), torch.log(mxy.clamp_min(1e-12) / j0y),
                     tune(zx, zpx), tune(zy, zpy), nal / n_turns,
                     torch.cos(th0), torch.sin(th0)], dim=1)
    return torch.nan_to_num(f, nan=0.0, posinf=50.0, neginf=-50.0).to(torch.float32)


def surrogate_da_scan(line, r_max, n_r, n_angles, num_turns, nemitt_x, nemitt_y,
                      deltas=(0.0,), feature_turns=256, train_fraction=0.15, band=1.5,
                      epochs=150, seed=0, net=None, feature_chunk=200_000):
    """DA scan screened by the surrogate; returns the survival turns used
    for the aperture (exact where tracked, predicted elsewhere), the
    per-angle apertures, which particles were tracked in full, timings and
    the network.  The line must have a device tracker.

    band: half-width, in beam sigmas, of the amplitude band around each
    angle's predicted aperture whose particles are tracked to num_turns."""
    torch = _t()
    from . import da
    from .monitor import ParticlesMonitor
    from .twiss import (_reference, compute_one_turn_matrix_finite_differences,
                        find_closed_orbit, get_linear_normal_form)
    dev = line.tracker.backend.device
    t_all = time.perf_counter()
    r, th, d = da.polar_grid(r_max, n_r, n_angles, deltas)
    xn, yn = r * np.cos(th), r * np.sin(th)
    p = da.build_normalized_particles(line, xn, 0.0, yn, 0.0, d, nemitt_x, nemitt_y)
    p.move(_context=dev)
    n = len(r)
    ref = _reference(line, None)
    bg = float(ref._np('beta0')[0] * ref._np('gamma0')[0])
    sx, sy = math.sqrt(nemitt_x / bg), math.sqrt(nemitt_y / bg)
    co = find_closed_orbit(line, ref, method='4d')
    R, _ = compute_one_turn_matrix_finite_differences(line, ref, co, kill_cavity=True)
    invW4 = get_linear_normal_form(R, only_4d_block=True)[1][:4, :4]
    # 1. feature turns for every particle, in chunks of particle ids (the
    # turn-by-turn record holds every field of every particle and turn)
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    order = torch.argsort(p._arrays['particle_id'])
    feat = torch.empty((n, 16), dtype=torch.float32, device=dev)
    for c0 in range(0, n, feature_chunk):
        c1 = min(n, c0 + feature_chunk)
        slots = order[c0:c1]
        sub = _take(p, slots)
        mon = ParticlesMonitor(start_at_turn=0, stop_at_turn=feature_turns,
                               particle_id_range=(c0, c1), auto_to_numpy=False, _context=dev)
        line.track(sub, num_turns=feature_turns, turn_by_turn_monitor=mon)
        feat[c0:c1] = features(mon, xn[c0:c1], yn[c0:c1], d[c0:c1], invW4, sx, sy,
                               feature_turns)
        for nn, vv in p._arrays.items():
            if vv.dtype == torch.uint32:
                vv.view(torch.int32)[slots] = sub._arrays[nn].view(torch.int32)
            else:
                vv[slots] = sub._arrays[nn]
        del mon, sub
    torch.cuda.synchronize(dev)
    t_feat = time.perf_counter() - t0
    surv = np.full(n, -1, dtype=np.int64)
    st = p._arrays['state'][order].cpu().numpy()
    at = p._arrays['at_turn'][order].cpu().numpy()
    surv[st <= 0] = at[st <= 0]
    alive = np.flatnonzero(st > 0)
    exact = st <= 0
    # 2. labels: a random subset of the survivors tracked to num_turns
    rng = np.random.default_rng(seed)
    n_train = max(16, int(train_fraction * len(alive))) if len(alive) else 0
    train = np.sort(rng.choice(alive, size=min(n_train, len(alive)), replace=False))

    def track_rest(ids):
        if len(ids) == 0:
            return np.zeros(0, dtype=np.int64)
        sub = _take(p, order[torch.as_tensor(ids, device=dev)])
        line.track(sub, num_turns=num_turns - feature_turns)
        s = sub.sorted_by_id()
        return np.where(s['state'] > 0, num_turns, s['at_turn'])

    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    surv[train] = track_rest(train)
    exact[train] = True
    torch.cuda.synchronize(dev)
    t_train_track = time.perf_counter() - t0
    # 3. fit log(survival) on the matrix cores
    t0 = time.perf_counter()
    net = net or MLP(device=dev, seed=seed)
    fx = feat[torch.as_tensor(train, device=dev)]
    fy = torch.as_tensor(np.log(surv[train].astype(np.float64)), dtype=torch.float32,
                         device=dev)
    hist = net.fit(fx, fy, epochs=epochs, batch=min(1024, max(64, len(train) // 4)),
                   seed=seed)
    pred = np.exp(np.minimum(net.predict(feat).cpu().numpy().astype(np.float64),
                             math.log(1e12)))
    torch.cuda.synchronize(dev)
    t_fit = time.perf_counter() - t0
    # 4. band around each angle's predicted aperture, tracked in full
    guess = np.where(exact, surv, np.minimum(pred, num_turns)).astype(np.float64)
    band_ids = []
    for a in np.unique(th):
        for dd in np.unique(d):
            sel = np.flatnonzero((th == a) & (d == dd))
            sel = sel[np.argsort(r[sel])]
            lost = np.flatnonzero(guess[sel] < 0.5 * num_turns)
            r_pred = r[sel][lost[0]] if len(lost) else r_max
            inb = sel[(np.abs(r[sel] - r_pred) <= band) & ~exact[sel]]
            band_ids.append(inb)
    band_ids = np.sort(np.concatenate(band_ids)) if band_ids else np.zeros(0, np.int64)
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    surv[band_ids] = track_rest(band_ids)
    exact[band_ids] = True
    torch.cuda.synchronize(dev)
    t_band = time.perf_counter() - t0
    # 5. the rest keeps its prediction (survives num_turns when predicted to)
    rest = ~exact
    surv[rest] = np.where(pred[rest] >= 0.5 * num_turns, num_turns,
                          np.minimum(pred[rest], num_turns - 1)).astype(np.int64)
    return {'r': r, 'theta': th, 'delta': d, 'survival': surv, 'exact': exact,
            'tracked_full': np.union1d(train, band_ids), 'pred': pred,
            'da': da.dynamic_aperture(r, th, d, surv, num_turns), 'loss_history': hist,
            'time': {'features': t_feat, 'train_tracking': t_train_track, 'fit': t_fit,
                     'band_tracking': t_band, 'total': time.perf_counter() - t_all},
            'net': net}
