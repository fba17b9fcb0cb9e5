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
