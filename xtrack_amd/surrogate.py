"""Survival-turn surrogate for dynamic-aperture scans (no counterpart in
the reference).

A small MLP (16 -> 256 -> 256 -> 256 -> 1, ReLU) predicts log(survival
turn) of a particle from its initial normalised coordinates and features of
its first turns (tracked by the plain tracker).  Its products run on the
matrix cores through xtm_gemm_bf16 (BF16 operands, FP32 accumulation),
forward and backward; the FP32 master weights, the bias and loss sums and
the Adam update are torch tensor operations on the same device.

The screened scan (``screened_da_scan``, over many lattices at once on the
batch tracker) tracks every grid particle for the feature turns with the
feature accumulators updated on the device, tracks a random subset of the
survivors to the full turn count for labels, fits an ensemble of networks
to them, and tracks to the full turn count only the particles whose
predicted survival lies within a factor of the target or whose amplitude
lies near the predicted aperture (the boundary band); particles the
networks place well inside or outside are assigned their prediction.
Plain full tracking of the same grid is the reference the scan's aperture
error is measured against.
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


def _small_sum(a):
    """Sum over the first axis in reductions of at most 128 terms each
    (rows in groups of 128, then the group sums), so that every reduction
    kernel works within one block and leaves no global partials: the same
    kernels run eagerly and inside a captured HIP graph."""
    torch = _t()
    n = a.shape[0]
    g = -(-n // 128)
    if g * 128 != n:
        a = torch.cat([a, a.new_zeros((g * 128 - n,) + tuple(a.shape[1:]))])
    part = a.reshape((g, 128) + tuple(a.shape[1:])).sum(dim=1)
    while part.shape[0] > 128:
        g2 = -(-part.shape[0] // 128)
        if g2 * 128 != part.shape[0]:
            part = torch.cat([part, part.new_zeros((g2 * 128 - part.shape[0],) +
                                                   tuple(part.shape[1:]))])
        part = part.reshape((g2, 128) + tuple(part.shape[1:])).sum(dim=1)
    return part.sum(dim=0)


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
            gb[li] = _small_sum(dz)
            if li > 0:
                dh, _ = gemm(dzb, to_bf16(self.W[li]), trans_b=True)   # dz W^T
                dz = dh * (zs[li - 1] > 0).to(dh.dtype)
        return gW, gb

    def adam_step(self, gW, gb, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        """One Adam update of the weights and biases in place.  The step
        count and the bias corrections live on the device, so that the update
        can be captured in a HIP graph and replayed."""
        torch = _t()
        if self._adam is None:
            self._adam = {'t': torch.zeros((), dtype=torch.float32, device=self.device),
                          'm': [torch.zeros_like(p) for p in self.W + self.b],
                          'v': [torch.zeros_like(p) for p in self.W + self.b]}
        st = self._adam
        st['t'].add_(1.0)
        bc1 = 1.0 - torch.pow(beta1, st['t'])
        bc2 = 1.0 - torch.pow(beta2, st['t'])
        params, grads = self.W + self.b, gW + gb
        for k, (p, g) in enumerate(zip(params, grads)):
            st['m'][k].mul_(beta1).add_(g, alpha=1 - beta1)
            st['v'][k].mul_(beta2).addcmul_(g, g, value=1 - beta2)
            mh = st['m'][k] / bc1
            vh = st['v'][k] / bc2
            p.sub_(lr * mh / (vh.sqrt() + eps))

    def _train_step(self, xb, yb, cb, lr):
        """Forward, backward and Adam on one minibatch; returns the sum of
        squared errors (device scalar).  cb: None or the censored flags."""
        torch = _t()
        out = self.forward(xb, keep=True)
        err = out - yb
        if cb is not None:
            err = torch.where(cb & (err > 0), 0.0, err)
        gW, gb = self.backward(2.0 * err / len(xb))
        self.adam_step(gW, gb, lr=lr)
        return _small_sum(err * err)

    def fit(self, x, y, epochs=200, batch=1024, lr=2e-3, seed=0, censored=None, graph=True):
        """Mean-squared error on standardised inputs and targets; x (N, 16),
        y (N,) torch FP32 on the device.  censored: None, or a bool (N,)
        tensor marking targets that are lower bounds (particles alive at the
        last tracked turn): their error counts only where the output lies
        below the target, so the network may place them beyond it.  With
        graph=True the step on a full minibatch is captured once in a HIP
        graph (torch.cuda.CUDAGraph) and replayed with each minibatch copied
        into its input buffers, the same kernels in the same order as the
        eager step (graph=False); a short last minibatch runs eagerly.
        Returns the loss per epoch."""
        torch = _t()
        self.x_mean = x.mean(dim=0)
        self.x_std = x.std(dim=0).clamp_min(1e-12)
        self.y_mean, self.y_std = float(y.mean()), float(y.std().clamp_min(1e-12))
        xn = ((x - self.x_mean) / self.x_std).contiguous()
        yn = ((y - self.y_mean) / self.y_std).contiguous()
        cens = None if censored is None else censored.to(self.device, torch.bool)
        g = torch.Generator(device='cpu').manual_seed(seed)
        hist = []
        n = len(xn)
        B = min(int(batch), n)
        gr = None
        sx = sy = sc = sloss = None
        for _ in range(epochs):
            perm = torch.randperm(n, generator=g).to(self.device)
            tot = torch.zeros((), dtype=torch.float32, device=self.device)
            for s in range(0, n, B):
                idx = perm[s:s + B]
                full = len(idx) == B
                if graph and full and gr is not None:
                    sx.copy_(xn[idx])
                    sy.copy_(yn[idx])
                    if sc is not None:
                        sc.copy_(cens[idx])
                    gr.replay()
                    tot += sloss
                    continue
                tot += self._train_step(xn[idx], yn[idx],
                                        None if cens is None else cens[idx], lr)
                if graph and full and gr is None:
                    # the eager step above allocated the optimiser state and
                    # the product workspace; capture the next one
                    sx = torch.empty_like(xn[:B])
                    sy = torch.empty_like(yn[:B])
                    sc = None if cens is None else torch.empty_like(cens[:B])
                    torch.cuda.synchronize(self.device)
                    gr = torch.cuda.CUDAGraph()
                    with torch.cuda.graph(gr):
                        sloss = self._train_step(sx, sy, sc, lr)
            hist.append(float(tot) / n)
        return hist

    def predict(self, x, batch=65536):
        torch = _t()
        xn = (x - self.x_mean) / self.x_std
        out = [self.forward(xn[s:s + batch]) for s in range(0, len(xn), batch)]
        return torch.cat(out) * self.y_std + self.y_mean


# ---------------------------------------------------------------------------
# dynamic-aperture scans over one or many lattices (error seeds), on the batch
# tracker, plain and screened by the surrogate

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


def _put(p, idx, sub):
    """Write the particles of `sub` back into slots `idx` of `p`."""
    torch = _t()
    for nn, vv in p._arrays.items():
        if vv.dtype == torch.uint32:
            vv.view(torch.int32)[idx] = sub._arrays[nn].view(torch.int32)
        else:
            vv[idx] = sub._arrays[nn]


def track_features(line_or_batch, particles, num_turns, inv_w4, co, sx, sy, starts=None,
                   acc=None, window=None):
    """Track `particles` (a device particle set) for num_turns whole turns
    with the feature accumulators updated on the device at the start of every
    turn, between one-turn tracking launches (xtm_track_line_features /
    xtm_batch_track_features);
    returns the accumulators, a (XTM_FEAT_N, capacity) float64 tensor on the
    device, row k = accumulator k of every particle slot.

    line_or_batch: a Line with a device tracker, or a BatchTracker together
    with `starts` (the variants' particle ranges in `particles`).  acc and
    window: accumulators to continue (one column per particle slot) and the
    recorded turn window (start, stop); by default fresh accumulators and
    the window of these num_turns turns."""
    torch = _t()
    lib = _native.load()
    n = particles._capacity
    if acc is None:
        acc = torch.zeros((_native.XTM_FEAT_N, n), dtype=torch.float64,
                          device=particles._arrays['x'].device)
    elif tuple(acc.shape) != (_native.XTM_FEAT_N, n) or not acc.is_contiguous():
        raise ValueError('acc must be a contiguous (XTM_FEAT_N, capacity) tensor')
    if window is None:
        t0 = int(particles._arrays['at_turn'].min().item()) if n else 0
        window = (t0, t0 + num_turns)
    fst = _native.make_features_struct(acc, n, int(window[0]), int(window[1]),
                                       np.asarray(inv_w4), co, (1.0 / sx, 1.0 / sy))
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


def features_from_acc(acc, x_norm0, y_norm0, delta, window):
    """The 16 network inputs per particle from the in-kernel accumulators
    of a feature window of `window` turns: r, theta, delta, the initial
    actions, mean, spread and maximum of the normalised actions over the
    recorded turns (relative to the initial ones), the mean phase advance
    per turn in each plane (tunes), the fraction of the turns recorded
    alive, and log10 of the tune diffusion in each plane (the difference of
    the mean phase advances of the window's second and first halves, as in
    frequency-map analysis)."""
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
    # advances recorded at turns >= start + ceil(window / 2): the second half
    n2 = (acc[0] - max((int(window) + 1) // 2, 1)).clamp_min(0.0)
    n1 = (acc[0] - 1.0 - n2).clamp_min(0.0)
    ph1x, ph1y = acc[7] - acc[15], acc[8] - acc[16]
    dqx = (acc[15] / n2.clamp_min(1.0) - ph1x / n1.clamp_min(1.0)).abs() / (2 * math.pi)
    dqy = (acc[16] / n2.clamp_min(1.0) - ph1y / n1.clamp_min(1.0)).abs() / (2 * math.pi)
    both = (n1 > 0) & (n2 > 0)
    ldqx = torch.where(both, torch.log10(dqx + 1e-12), torch.zeros_like(dqx))
    ldqy = torch.where(both, torch.log10(dqy + 1e-12), torch.zeros_like(dqy))
    r0 = torch.sqrt(torch.as_tensor(x_norm0 ** 2 + y_norm0 ** 2, device=dev))
    th0 = torch.atan2(torch.as_tensor(y_norm0, device=dev), torch.as_tensor(x_norm0, device=dev))
    f = torch.stack([r0, th0, torch.as_tensor(delta * 1e3, device=dev),
                     torch.log1p(j0x), torch.log1p(j0y),
                     torch.log(mjx.clamp_min(1e-12) / j0x), torch.log(mjy.clamp_min(1e-12) / j0y),
                     sjx / mjx.clamp_min(1e-12), sjy / mjy.clamp_min(1e-12),
                     torch.log(acc[3].clamp_min(1e-12) / j0x),
                     torch.log(acc[6].clamp_min(1e-12) / j0y),
                     qx, qy, acc[0] / nturn, ldqx, ldqy], dim=1)
    return torch.nan_to_num(f, nan=0.0, posinf=50.0, neginf=-50.0).to(torch.float32)


class GridSet:
    """The grid particles of every lattice in one device particle set:
    lattice l holds slots starts[l]..starts[l+1] on its own polar grid of
    n_r[l] amplitudes (0-r_max beam sigmas) x n_angles angles x deltas
    (da.polar_grid order), particle_id = slot.  Per slot: r, theta and
    delta of its grid point (r_slot, theta_slot, delta_slot)."""

    def __init__(self, lines, r_max, n_r, n_angles, deltas, nemitt_x, nemitt_y, device):
        from . import da
        from .particles import Particles
        from .twiss import _reference
        self.n_lines = len(lines)
        n_r = [int(n_r)] * self.n_lines if np.ndim(n_r) == 0 else [int(v) for v in n_r]
        if len(n_r) != self.n_lines:
            raise ValueError('one amplitude count per lattice is needed')
        self.r_max = float(r_max)
        hosts, rs, ths, ds = [], [], [], []
        for ln, nr in zip(lines, n_r):
            r, th, d = da.polar_grid(r_max, nr, n_angles, deltas)
            pp = da.build_normalized_particles(ln, r * np.cos(th), 0.0, r * np.sin(th), 0.0, d,
                                               nemitt_x, nemitt_y)
            hosts.append(pp.to_numpy())
            rs.append(r)
            ths.append(th)
            ds.append(d)
        self.r_slot = np.concatenate(rs)
        self.theta_slot = np.concatenate(ths)
        self.delta_slot = np.concatenate(ds)
        sizes = [len(r) for r in rs]
        self.starts = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
        self.n_slots = int(self.starts[-1])
        p0 = _reference(lines[0], None)
        big = Particles.__new__(Particles)
        big._capacity = self.n_slots
        big.q0, big.mass0, big.t_sim = p0.q0, p0.mass0, 0.0
        big.start_tracking_at_element = -1
        big._arrays = {nn: np.concatenate([h[nn] for h in hosts]) for nn in hosts[0]}
        big._arrays['particle_id'] = np.arange(self.n_slots, dtype=np.int64)
        big._device = None
        big.move(_context=device)
        self.particles = big
        self._initial = {nn: vv.clone() for nn, vv in big._arrays.items()}
        self.x_norm = self.r_slot * np.cos(self.theta_slot)
        self.y_norm = self.r_slot * np.sin(self.theta_slot)
        bg = float(p0._np('beta0')[0] * p0._np('gamma0')[0])
        self.sigma = (math.sqrt(nemitt_x / bg), math.sqrt(nemitt_y / bg))
        self.device = device

    def line_of(self, slots):
        return np.searchsorted(self.starts, slots, side='right') - 1

    def reset(self):
        """Back to the initial grid particles (for another scan)."""
        for nn, vv in self.particles._arrays.items():
            vv.copy_(self._initial[nn])


def _starts_of(slots_sorted, starts):
    """Layout of a sorted subset of slots: the start of each lattice's
    range within the subset."""
    return np.searchsorted(slots_sorted, starts).astype(np.int64)


def track_subset(bt, gs, slots, num_turns, block=None, features=None, progress=None,
                 prune=None):
    """Track the grid particles `slots` (sorted numpy slot indices) of
    GridSet gs for num_turns turns on BatchTracker bt, compacting the lost
    ones out every `block` turns (all turns in one launch when None), and
    write the results back into gs.particles.  features: None, or
    (inv_w4, co) to gather the feature accumulators (with block None);
    returns them for the slots (columns in the order of `slots`, over the
    window of these num_turns turns, continued across blocks), else None.
    progress: None, or called as progress(turns done, particles still
    tracked) after each block.  prune: None, or a function of the slots
    still alive after a block that returns those to keep tracking (the
    others stop where they are)."""
    torch = _t()
    dev = gs.device
    slots = np.asarray(slots, dtype=np.int64)
    done, acc = 0, None
    if features is not None:
        acc = torch.zeros((_native.XTM_FEAT_N, len(slots)), dtype=torch.float64, device=dev)
        col0 = slots.copy()
        t0 = int(gs.particles._arrays['at_turn'][torch.as_tensor(slots, device=dev)]
                 .min().item()) if len(slots) else 0
    while done < num_turns and len(slots):
        n = num_turns - done if block is None else min(block, num_turns - done)
        idx = torch.as_tensor(slots, device=dev)
        sub = _take(gs.particles, idx)
        lay = _starts_of(slots, gs.starts)
        if features is not None:
            inv_w4, co = features
            cols = torch.as_tensor(np.searchsorted(col0, slots), device=dev)
            acc_sub = acc[:, cols].contiguous()
            track_features(bt, sub, n, inv_w4, co, gs.sigma[0], gs.sigma[1], starts=lay,
                           acc=acc_sub, window=(t0, t0 + num_turns))
            acc[:, cols] = acc_sub
        else:
            bt._set_layout(lay, bt.config or (256, 4))
            bt._launch(sub, n)
        _put(gs.particles, idx, sub)
        done += n
        if done < num_turns:
            alive = (sub._arrays['state'] > 0).cpu().numpy()
            slots = slots[alive]
            if prune is not None and len(slots):
                slots = prune(slots)
        if progress is not None:
            progress(done, len(slots))
    torch.cuda.synchronize(dev)
    return acc


def survival_of(gs, num_turns):
    """Per slot: the turn of loss, or num_turns for a particle alive."""
    st = gs.particles._arrays['state'].cpu().numpy()
    at = gs.particles._arrays['at_turn'].cpu().numpy()
    return np.where(st > 0, num_turns, at).astype(np.int64)


def da_per_line(gs, surv, num_turns):
    """[{(theta, delta): DA}] per lattice (da.dynamic_aperture)."""
    from . import da
    out = []
    for l in range(gs.n_lines):
        sl = slice(gs.starts[l], gs.starts[l + 1])
        out.append(da.dynamic_aperture(gs.r_slot[sl], gs.theta_slot[sl], gs.delta_slot[sl],
                                       surv[sl], num_turns))
    return out


def plain_da_scan_batch(bt, gs, num_turns, block=256, progress=None):
    """Every grid particle of every lattice tracked to num_turns (lost
    particles compacted out every `block` turns).  Returns survival turns,
    the apertures per lattice and the wall time."""
    torch = _t()
    torch.cuda.synchronize(gs.device)
    t0 = time.perf_counter()
    track_subset(bt, gs, np.arange(gs.n_slots), num_turns, block=block, progress=progress)
    wall = time.perf_counter() - t0
    surv = survival_of(gs, num_turns)
    return {'survival': surv, 'da': da_per_line(gs, surv, num_turns), 'time': wall}


def _group_keys(gs, slots):
    """(lattice, angle index, delta index) of each slot."""
    th_idx = np.searchsorted(np.unique(gs.theta_slot), gs.theta_slot[slots])
    d_idx = np.searchsorted(np.unique(gs.delta_slot), gs.delta_slot[slots])
    return gs.line_of(slots), th_idx, d_idx


def interp_predictor(gs, labelled, surv, candidates):
    """Survival prediction without a network, for comparison: within each
    (lattice, angle, delta) group, log survival interpolated in amplitude
    between the labelled particles of the group (their end values beyond
    them; 1 turn for a group without labels)."""
    lab = np.asarray(labelled, dtype=np.int64)
    cand = np.asarray(candidates, dtype=np.int64)
    pred = np.ones(len(cand))
    lk = np.stack(_group_keys(gs, lab), axis=1) if len(lab) else np.zeros((0, 3), np.int64)
    ck = np.stack(_group_keys(gs, cand), axis=1)
    groups = {}
    for i, k in enumerate(map(tuple, lk)):
        groups.setdefault(k, []).append(lab[i])
    by_key = {}
    for i, k in enumerate(map(tuple, ck)):
        by_key.setdefault(k, []).append(i)
    for k, members in by_key.items():
        g = groups.get(k)
        if not g:
            continue
        g = np.asarray(g)
        g = g[np.argsort(gs.r_slot[g])]
        m = np.asarray(members)
        pred[m] = np.exp(np.interp(gs.r_slot[cand[m]], gs.r_slot[g],
                                   np.log(np.maximum(surv[g], 1))))
    return pred


def _group_id(gs):
    """Per slot: the index of its (lattice, angle, delta) group."""
    if getattr(gs, '_gid', None) is None:
        keys = np.stack(_group_keys(gs, np.arange(gs.n_slots)), axis=1)
        gs._gid = np.unique(keys, axis=0, return_inverse=True)[1].reshape(-1)
    return gs._gid


def below_first_loss(gs, lost):
    """Per slot: True when its amplitude lies below the smallest amplitude
    of a particle of `lost` (bool per slot) in its (lattice, angle, delta)
    group.  Only these slots can still change the group's aperture (the
    largest amplitude below which every particle survived)."""
    gid = _group_id(gs)
    first = np.full(int(gid.max()) + 1 if len(gid) else 0, np.inf)
    np.minimum.at(first, gid[lost], gs.r_slot[lost])
    return gs.r_slot < first[gid]


def boundary_band(gs, guess, rest, num_turns, band_factor, band_sigma, lost_below=0.5,
                  guess_lo=None, guess_hi=None):
    """The slots of `rest` to track in full.  guess: survival per slot,
    exact where known, predicted for the slots of `rest`; guess_lo and
    guess_hi: the smallest and largest prediction of an ensemble (default
    guess).  A slot of `rest` is in the band when its predicted survival
    range meets [num_turns / band_factor, num_turns * band_factor], or its
    amplitude lies within band_sigma of the predicted aperture of its
    (lattice, angle, delta) group: the smallest amplitude guessed lost
    (known survival below num_turns, or smallest prediction below
    lost_below * num_turns; the largest amplitude when none is)."""
    rest = np.asarray(rest, dtype=np.int64)
    lo_all = guess if guess_lo is None else guess_lo
    hi_all = guess if guess_hi is None else guess_hi
    lo, hi = lo_all[rest], hi_all[rest]
    in_band = (hi >= num_turns / band_factor) & (lo <= num_turns * band_factor)
    n_all = len(guess)
    r_all = gs.r_slot
    keys = np.stack(_group_keys(gs, np.arange(n_all)), axis=1)
    gid = np.unique(keys, axis=0, return_inverse=True)[1].reshape(-1)
    lost = guess < num_turns
    lost[rest] = lo < lost_below * num_turns
    lost_r = np.full(int(gid.max()) + 1, np.inf)
    np.minimum.at(lost_r, gid[lost], r_all[lost])
    r_pred = np.where(np.isfinite(lost_r), lost_r, gs.r_max)
    in_band |= np.abs(r_all[rest] - r_pred[gid[rest]]) <= band_sigma
    return rest[in_band]


def screened_da_scan(bt, gs, num_turns, feature_turns=256, label_fraction=0.02,
                     band_factor=3.0, band_sigma=0.5, epochs=150, seed=0, net=None,
                     predictor='mlp', block=256, normalisation=None, censor_margin=None,
                     ensemble=4, prune_band=True, progress=None):
    """Dynamic-aperture scan of every lattice of GridSet gs screened by the
    survival surrogate, on BatchTracker bt (the lattices of gs, in order):

    1. every grid particle is tracked for feature_turns turns with the
       feature accumulators gathered in the kernel (with prune_band, a
       particle stops at the next 8-turn compaction once a particle below
       it in its group is lost: it can no longer change the aperture and is
       neither labelled nor banded; its survival is then predicted from
       the features of the turns it made);
    2. a random label_fraction of the survivors of each lattice that lie
       below the first loss of their (angle, delta) group (only those can
       change the aperture) is tracked to num_turns: their survival turns
       are the labels;
    3. the network (16-256-256-256-1, BF16 products on the matrix cores) is
       fit to log survival turn on the labels; for a label alive at
       num_turns the target is a lower bound, censor_margin x num_turns
       (default 2), counted only where the output lies below it, so that
       the network may place stable particles beyond the band; `ensemble`
       networks from different initialisations and minibatch orders are
       fit (predictor='interp': no network, per-angle interpolation of the
       labels in amplitude);
    4. of the untracked survivors below the first known loss of their
       group, those for which any network predicts a survival within a
       factor band_factor of num_turns, or whose amplitude lies within
       band_sigma beam sigmas of their angle's predicted aperture (the
       smallest amplitude that any network predicts lost), are tracked to
       num_turns (the boundary band); with prune_band, a band particle
       that finds itself above a loss of its group (one of the band lost
       earlier in the pass) stops at the next compaction, since it can no
       longer change the aperture;
    5. every other particle keeps its prediction (the ensemble's median):
       alive if predicted to survive at least half of num_turns, else lost
       at its predicted turn.

    Tracking is continued, not restarted: a particle tracked in full has
    exactly its plain-scan survival turn.  normalisation: (inv_w4, co) of
    the features, default the first lattice's 4d normal form at delta 0."""
    torch = _t()
    dev = gs.device
    t_all = time.perf_counter()
    n_all = gs.n_slots
    if normalisation is None:
        normalisation = normal_form_4d(bt.lines[0], dev)
    times = {}
    # 1. feature turns
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    # a particle above a loss of its (lattice, angle, delta) group can no
    # longer change the group's aperture: it stops at the next compaction
    # (the feature pass compacts every 8 turns, the band every `block`)
    def _above_no_loss(slots):
        lost_now = gs.particles._arrays['state'].cpu().numpy() <= 0
        return slots[below_first_loss(gs, lost_now)[slots]]

    acc = track_subset(bt, gs, np.arange(n_all), feature_turns, features=normalisation,
                       block=8 if prune_band else None,
                       prune=_above_no_loss if prune_band else None)
    feat = features_from_acc(acc, gs.x_norm, gs.y_norm, gs.delta_slot, feature_turns)
    torch.cuda.synchronize(dev)
    times['features'] = time.perf_counter() - t0
    surv = survival_of(gs, feature_turns)
    exact = surv < feature_turns
    # particles above a loss of their group cannot change its aperture: the
    # labels and the band are drawn from the survivors below the first loss
    alive = np.flatnonzero(~exact & below_first_loss(gs, exact))
    # 2. labels
    rng = np.random.default_rng(seed)
    lab = []
    for l in range(gs.n_lines):
        al = alive[(alive >= gs.starts[l]) & (alive < gs.starts[l + 1])]
        k = min(len(al), max(16, int(round(label_fraction * len(al)))))
        if k:
            lab.append(np.sort(rng.choice(al, size=k, replace=False)))
    lab = np.concatenate(lab) if lab else np.zeros(0, np.int64)
    t0 = time.perf_counter()
    track_subset(bt, gs, lab, num_turns - feature_turns, block=block,
                 progress=None if progress is None else
                 (lambda d, n: progress('labels', feature_turns + d, n)))
    times['label_tracking'] = time.perf_counter() - t0
    surv = survival_of(gs, num_turns)
    exact[lab] = True
    # 3. predictor
    t0 = time.perf_counter()
    rest = np.flatnonzero(~exact)
    hist = None
    nets = []
    if predictor == 'mlp':
        fx = feat[torch.as_tensor(lab, device=dev)]
        alive_lab = surv[lab] >= num_turns
        margin = 2.0 if censor_margin is None else float(censor_margin)
        fy = torch.as_tensor(np.log(np.where(alive_lab, margin * num_turns,
                                             surv[lab].astype(np.float64))),
                             dtype=torch.float32, device=dev)
        cens = torch.as_tensor(alive_lab, device=dev)
        fr = feat[torch.as_tensor(rest, device=dev)]
        preds = []
        for k in range(int(ensemble)):
            nk = net if (k == 0 and net is not None) else MLP(device=dev, seed=seed + k)
            h = nk.fit(fx, fy, epochs=epochs, batch=min(1024, max(64, len(lab) // 4)),
                       seed=seed + k, censored=cens)
            hist = h if hist is None else hist
            preds.append(np.exp(np.minimum(nk.predict(fr).cpu().numpy().astype(np.float64),
                                           math.log(1e12))))
            nets.append(nk)
        P = np.stack(preds)
        pred, p_lo, p_hi = np.median(P, axis=0), P.min(axis=0), P.max(axis=0)
    elif predictor == 'interp':
        pred = interp_predictor(gs, lab, surv, rest)
        p_lo = p_hi = pred
    else:
        raise ValueError("predictor must be 'mlp' or 'interp'")
    torch.cuda.synchronize(dev)
    times['fit_predict'] = time.perf_counter() - t0
    # 4. boundary band: predicted survival near num_turns, or amplitude near
    # the predicted aperture of its (lattice, angle, delta) group
    guess = surv.astype(np.float64)
    guess[rest] = pred
    lo, hi = guess.copy(), guess.copy()
    lo[rest], hi[rest] = p_lo, p_hi
    relevant = rest[below_first_loss(gs, exact & (surv < num_turns))[rest]]
    band = boundary_band(gs, guess, relevant, num_turns, band_factor, band_sigma,
                         guess_lo=lo, guess_hi=hi)
    t0 = time.perf_counter()
    track_subset(bt, gs, band, num_turns - feature_turns, block=block,
                 progress=None if progress is None else
                 (lambda d, n: progress('band', feature_turns + d, n)),
                 prune=_above_no_loss if prune_band else None)
    times['band_tracking'] = time.perf_counter() - t0
    surv_t = survival_of(gs, num_turns)
    reached = gs.particles._arrays['at_turn'].cpu().numpy()
    st_now = gs.particles._arrays['state'].cpu().numpy()
    full = band[(st_now[band] <= 0) | (reached[band] >= num_turns)]
    surv[full] = surv_t[full]
    exact[full] = True
    # 5. the rest keeps its prediction (the ensemble's median): alive when
    # predicted to survive half the turns or more
    # (a band particle stopped early keeps at least the turn it reached)
    left = np.flatnonzero(~exact)
    p_left = guess[left]
    low = np.maximum(feature_turns, np.minimum(reached[left], num_turns - 1))
    surv[left] = np.where(p_left >= 0.5 * num_turns, num_turns,
                          np.clip(p_left, low, num_turns - 1)).astype(np.int64)
    times['total'] = time.perf_counter() - t_all
    return {'survival': surv, 'exact': exact, 'labelled': lab, 'band': band,
            'band_stopped_above_a_loss': int(len(band) - len(full)),
            'tracked_full_fraction': (len(lab) + len(full)) / n_all,
            'da': da_per_line(gs, surv, num_turns), 'loss_history': hist, 'time': times,
            'net': nets[0] if nets else None, 'nets': nets}


def normal_form_4d(line, device):
    """(inv_w4, closed orbit x, px, y, py) of the line's 4d linear normal
    form at delta 0: the normalisation of the features."""
    from .twiss import (_reference, compute_one_turn_matrix_finite_differences,
                        find_closed_orbit, get_linear_normal_form)
    if line.tracker is None:
        line.build_tracker(device=device)
    ref = _reference(line, None)
    co = find_closed_orbit(line, ref, method='4d')
    R, _ = compute_one_turn_matrix_finite_differences(line, ref, co, kill_cavity=True)
    inv_w4 = get_linear_normal_form(R, only_4d_block=True)[1][:4, :4]
    return inv_w4, [float(v) for v in co[:4]]


def da_errors(da_a, da_b):
    """Per lattice: the largest |difference| of the per-angle apertures,
    the difference of their means and the difference of their minima (the
    lattice's aperture over all angles and deltas), in beam sigmas."""
    out = []
    for a, b in zip(da_a, da_b):
        d = np.array([a[k] - b[k] for k in b])
        out.append({'max_abs': float(np.max(np.abs(d))),
                    'mean_diff': float(np.mean([a[k] for k in b]) - np.mean(list(b.values()))),
                    'min_diff': float(min(a[k] for k in b) - min(b.values()))})
    return out


def surrogate_da_scan(line, r_max, n_r, n_angles, num_turns, nemitt_x, nemitt_y,
                      deltas=(0.0,), feature_turns=256, train_fraction=0.15, band=1.5,
                      epochs=150, seed=0, net=None):
    """Screened scan of one line (screened_da_scan with the band in beam
    sigmas around each angle's predicted aperture only); returns the grid,
    the survival turns (exact where tracked, predicted elsewhere) in grid
    order, the per-angle apertures, which particles were tracked in full,
    timings and the network.  The line must have a device tracker."""
    from .batch import BatchTracker
    dev = line.tracker.backend.device
    bt = BatchTracker([line], device=dev, compile='auto')
    gs = GridSet([line], r_max, n_r, n_angles, deltas, nemitt_x, nemitt_y, dev)
    res = screened_da_scan(bt, gs, num_turns, feature_turns=feature_turns,
                           label_fraction=train_fraction, band_factor=1.0, band_sigma=band,
                           epochs=epochs, seed=seed, net=net, ensemble=1)
    return {'r': gs.r_slot, 'theta': gs.theta_slot, 'delta': gs.delta_slot,
            'survival': res['survival'],
            'exact': res['exact'], 'tracked_full': np.union1d(res['labelled'], res['band']),
            'da': res['da'][0], 'loss_history': res['loss_history'], 'time': res['time'],
            'net': res['net']}
