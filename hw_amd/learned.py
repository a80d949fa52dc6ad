"""Training loop of the learned EM initialiser (no reference counterpart).

The plain pipeline generates the training pairs: an analyzer with capture
armed writes, at every plain EM estimate, each ZMW's network input (its
histogram re-binned to 128 bins) and its target (the converged EM result in
the network's output units) with a validity flag.  Training runs the
library's own GPU forward/backward/Adam kernels over device-resident pairs.
"""
import math
import time

import numpy as np

from . import (BatchAnalyzer, DevBuf, MLP_IN, MLP_OUT, bench_config, check, dev_bytes, lib, ptr, sync,
               tracegen, tracegen_default)
from . import types as T

import ctypes


class PairSet:
    """Device-resident (input, target, valid) pairs."""

    def __init__(self, capacity):
        self.capacity = int(capacity)
        self.n = 0
        self.x = dev_bytes(self.capacity * MLP_IN * 2)
        self.t = dev_bytes(self.capacity * MLP_OUT * 4)
        self.v = dev_bytes(self.capacity)

    def append(self, d_x, d_t, d_v, count):
        count = min(count, self.capacity - self.n)
        if count <= 0:
            return 0
        for dst, src, w in ((self.x, d_x, MLP_IN * 2), (self.t, d_t, MLP_OUT * 4), (self.v, d_v, 1)):
            check(lib.hwpa_memcpy_dtod(ctypes.c_void_p(dst.data_ptr() + self.n * w), ptr(src),
                                       ctypes.c_size_t(count * w)))
        self.n += count
        return count

    def valid_fraction(self):
        v = self.v.download(np.uint8, (self.capacity,))[:self.n]
        return float(v.mean()) if self.n else 0.0


def generate_pairs(lanes, chunks, seed, snr=(4.0, 20.0), frames=512, cfg=None, pairs=None, min_conf=0.1,
                   gen_overrides=None, log=None, limit=None):
    """Runs the plain pipeline on `chunks` chunks of seeded synthetic traces with
    per-ZMW SNR ~ U[snr] and collects one pair per ZMW and estimate (at most
    `limit` pairs from this call)."""
    cfg = cfg if cfg is not None else bench_config(frames)
    kw = dict(seed=seed, snr_lo=float(snr[0]), snr_hi=float(snr[1]))
    kw.update(gen_overrides or {})
    gen = tracegen_default(**kw)
    an = BatchAnalyzer(0, lanes, frames, cfg)
    Z = lanes * T.LANE
    cx, ct, cv = dev_bytes(Z * MLP_IN * 2), dev_bytes(Z * MLP_OUT * 4), dev_bytes(Z)
    an.capture_training(cx, ct, cv, min_conf)
    if pairs is None:
        pairs = PairSet(Z * chunks)
    d_state = dev_bytes(lanes * T.TRACEGEN_STATE.itemsize)
    d_tr = dev_bytes(lanes * frames * T.LANE * 2)
    t0 = time.perf_counter()
    start = pairs.n
    for c in range(chunks):
        if limit is not None and pairs.n - start >= limit:
            break
        tracegen(gen, lanes, frames, c * frames, d_state, c == 0, d_tr)
        an(d_tr, c * frames)
        if an.flags()[1]:
            an.sync()
            pairs.append(cx, ct, cv, Z if limit is None else min(Z, limit - (pairs.n - start)))
    sync()
    if log:
        log("generated %d pairs (%d lanes x %d chunks) in %.1f s" % (pairs.n, lanes, chunks, time.perf_counter() - t0))
    an.capture_training(None, None, None)
    return pairs


# Loss weights per squared raw unit (u = 2 binSize = 0.4 sigma for sigma >= 5):
# the approximate EM log-likelihood lost by a start that is off by one unit, for
# a 512-frame histogram with ~360 baseline frames and analog modes ~2 sigma
# wide -- baseline mean N_bl (0.4)^2 / 2, log variance N_bl / 4, analog mean
# N_a (0.2)^2 / 2 -- and N x KL for the mixture weights, all divided by N.
LL_WEIGHTS = (29.0, 90.0, 0.74, 0.74, 0.74, 0.74)
LL_FRAMES = 512.0


def fit_output_map(init, pairs):
    """Sets the network's output map from the valid training targets: offset
    and scale standardise each raw target, weights follow LL_WEIGHTS."""
    n = pairs.n
    t = pairs.t.download(np.float32, (pairs.capacity, MLP_OUT))[:n]
    v = pairs.v.download(np.uint8, (pairs.capacity,))[:n].astype(bool)
    tv = t[v][:, :6].astype(np.float64)
    off = tv.mean(axis=0)
    sc = np.maximum(tv.std(axis=0), 1e-3)
    w = [LL_WEIGHTS[j] * sc[j] ** 2 / LL_FRAMES for j in range(6)] + [1.0, 1.0]
    init.set_output_map(off, sc, w)
    return off, sc, w


def train(init, pairs, epochs=4, batch=4096, lr=2e-3, lr_end=1e-4, log=None, fit_map=True, seed=5):
    """Adam over the pairs in batches of consecutive pairs, the batches of each
    epoch in a seeded random order, cosine-decayed learning rate; returns the
    loss of every 50th step (and the last)."""
    n_batches = pairs.n // batch
    if n_batches == 0:
        raise ValueError("fewer pairs than one batch")
    if fit_map:
        fit_output_map(init, pairs)
    steps = epochs * n_batches
    d_loss = dev_bytes(4)
    losses = []
    rng = np.random.default_rng(seed)
    order = np.concatenate([rng.permutation(n_batches) for _ in range(epochs)])
    t0 = time.perf_counter()
    for s in range(steps):
        b = int(order[s])
        f = 0.5 * (1 + math.cos(math.pi * s / max(steps - 1, 1)))
        rate = lr_end + (lr - lr_end) * f
        init.train_step(pairs.x.view(b * batch * MLP_IN * 2, batch * MLP_IN * 2),
                        pairs.t.view(b * batch * MLP_OUT * 4, batch * MLP_OUT * 4), batch, rate, d_loss=d_loss,
                        d_valid=pairs.v.view(b * batch, batch))
        if s % 50 == 0 or s == steps - 1:
            sync()
            losses.append((s, float(d_loss.download(np.float32, (1,))[0])))
    sync()
    dt = time.perf_counter() - t0
    if log:
        log("trained %d steps (batch %d, %d epochs) in %.1f s: loss %.4f -> %.4f"
            % (steps, batch, epochs, dt, losses[0][1], losses[-1][1]))
    return losses, dt
