"""The histogram kernels' divide-free bin (hw_shortbin, include/hwpa_math.h)
against the reference's floorCastInt((x - lo) / binSize) with the IEEE divide
(TraceHistogramAccumHost::AddBlock, UHistogramSimd::BinIndex), for every int16
sample: it may decline a sample (the kernel then takes the divide) but must
never decide one wrongly, nor decide one whose bin lies outside the histogram.
The bounds include what the reference's reset produces from baseline
statistics (binSize = max(0.2 sigma, 1), lo = mu - 4 sigma, binSize recomputed
as (hi - lo) / 300), integral and round bounds (quotients exactly on bin edges),
bounds whose quotients come within a few ulps of the edges, and degenerate
ones (zero, negative, NaN, tiny and huge bin sizes, infinite bounds)."""
import ctypes

import numpy as np


def _sweep(O, lo, bs):
    lo = np.ascontiguousarray(lo, dtype=np.float32)
    bs = np.ascontiguousarray(bs, dtype=np.float32)
    out = (ctypes.c_uint64 * 3)()
    O.lib.orc_shortbin_sweep(lo.ctypes.data_as(ctypes.c_void_p), bs.ctypes.data_as(ctypes.c_void_p),
                             ctypes.c_int32(len(lo)), out)
    return out[0], out[1], out[2]


def _reset_bounds(mu, sigma):
    """TraceHistogramAccumHost::ResetImpl's bounds in float32 arithmetic."""
    f = np.float32
    bs0 = np.maximum(f(0.2) * sigma, f(1.0)).astype(np.float32)
    lo = (mu - f(4.0) * sigma).astype(np.float32)
    hi = (lo + f(300.0) * bs0).astype(np.float32)
    return lo, ((hi - lo) / f(300.0)).astype(np.float32)


def test_shortbin_reference_reset_bounds(O):
    rng = np.random.default_rng(11)
    mu = rng.uniform(-30, 30, 400).astype(np.float32)
    sigma = rng.uniform(2, 40, 400).astype(np.float32)
    sigma[:40] = 10.0                                     # the fallback sigma with mean 0
    mu[:40] = 0.0
    lo, bs = _reset_bounds(mu, sigma)
    in_range, decided, wrong = _sweep(O, lo, bs)
    assert wrong == 0
    assert in_range > 0 and decided / in_range > 0.999


def test_shortbin_round_and_near_edge_bounds(O):
    los, bss = [], []
    for lo in (-40.0, -48.0, 0.0, -1.0, -100.5, -47.99999, 3.0, -1e-20, 1e-20, -0.3):
        for bs in (2.0, 2.4, 3.3, 1.7, 2.43, 0.3, 1.1, 7.9, 0.1, 0.2, 0.7, 1.0, 3.0, 0.25, 1e-3, 5.0, 2.5, 0.6, 1.2):
            los.append(lo)
            bss.append(bs)
    # bin sizes one ulp either side of round values: quotients a few ulps from the edges
    for bs in (0.2, 0.3, 2.4, 1.0, 3.0):
        b = np.float32(bs)
        for nb in (np.nextafter(b, np.float32(0)), np.nextafter(b, np.float32(10))):
            los.append(-40.0)
            bss.append(float(nb))
    in_range, decided, wrong = _sweep(O, los, bss)
    assert wrong == 0
    # inexact bin sizes (2.4, 0.3, ...) put integral quotients inside the rounding band: those
    # samples take the divide
    assert decided / in_range > 0.9


def test_shortbin_degenerate_bounds_decline(O):
    nan, inf = float("nan"), float("inf")
    los = [-40.0, -40.0, -40.0, -40.0, -40.0, -40.0, nan, inf, -inf, -1e9, 1e9, -40.0]
    bss = [0.0, -2.0, nan, inf, 1e-30, 3e30, 2.0, 2.0, 2.0, 2.0, 2.0, 1e-40]
    in_range, decided, wrong = _sweep(O, los, bss)
    assert wrong == 0
    for lo, bs in zip(los, bss):
        if not (bs > 0) or bs == inf or bs < 1e-38 or lo != lo or abs(lo) == inf:
            assert _sweep(O, [lo], [bs])[1] == 0, (lo, bs)
