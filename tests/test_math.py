"""The deterministic math of include/hwpa_math.h (shared by the kernels and
the oracle) against libm/numpy: ulp-level accuracy, exact special cases, and
IEEE half conversion matching numpy's float16."""
import ctypes

import numpy as np
import pytest


def _eval(O, kind, a, b=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = np.ascontiguousarray(b if b is not None else np.zeros_like(a), dtype=np.float32)
    out = np.empty_like(a)
    O.lib.orc_math(ctypes.c_int32(kind), a.ctypes.data_as(ctypes.c_void_p), b.ctypes.data_as(ctypes.c_void_p),
                   out.ctypes.data_as(ctypes.c_void_p), ctypes.c_int32(a.size))
    return out


def _ulps(x, ref):
    x = x.astype(np.float64)
    ref = ref.astype(np.float64)
    spacing = np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64)
    return np.abs(x - ref) / spacing


def test_exp_accuracy(O):
    x = np.concatenate([np.linspace(-87, 88, 200001, dtype=np.float32),
                        np.float32([0.0, -0.0, 1e-30, -1e-30, 0.5, -0.5])])
    got = _eval(O, 0, x)
    ref = np.exp(x.astype(np.float64))
    assert _ulps(got, ref).max() <= 3.0
    assert _eval(O, 0, np.float32([0.0]))[0] == 1.0


def test_exp_limits(O):
    got = _eval(O, 0, np.float32([100.0, -110.0, np.inf, -np.inf, np.nan, -100.0]))
    assert got[0] == np.inf and got[1] == 0.0 and got[2] == np.inf and got[3] == 0.0 and np.isnan(got[4])
    # subnormal results keep their value to within a few subnormal steps
    assert abs(got[5] - np.float32(np.exp(-100.0))) <= 4 * np.float32(1.4e-45)


def test_exp_is_positive_zero_below_minus_105(O):
    """The EM kernel skips exps of boundaries whose exponents all lie below
    -105 and uses +0 instead; that is exact only if exp returns +0 there.
    Every float in [-106, -105) and a log-spaced sweep down to -FLT_MAX."""
    lo = np.float32(-106.0).view(np.uint32)
    hi = np.float32(-105.0).view(np.uint32)
    dense = np.arange(hi, lo + 1, dtype=np.uint32).view(np.float32)   # negative floats: bits grow with |x|
    sweep = -np.logspace(np.log10(105.0), np.log10(3.0e38), 100001).astype(np.float32)
    x = np.concatenate([dense, sweep, np.float32([-np.inf, -3.4028235e38])])
    got = _eval(O, 0, x)
    assert np.all(got == 0.0) and not np.signbit(got).any()
    # and -104 is still above the threshold of a non-zero result being possible
    assert _eval(O, 0, np.float32([-103.0]))[0] > 0.0


def test_log_accuracy(O):
    x = np.concatenate([np.exp(np.linspace(-80, 80, 200001)).astype(np.float32),
                        np.float32([1.0, 2.0, 0.5, 1.5e-42, 3.4e38, 1.0000001, 0.99999994])])
    got = _eval(O, 1, x)
    ref = np.log(x.astype(np.float64))
    err = np.abs(got.astype(np.float64) - ref)
    tol = np.maximum(3 * np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64), 2e-7 * np.abs(ref) + 1e-12)
    assert (err <= np.maximum(tol, 3e-8)).all()
    assert _eval(O, 1, np.float32([1.0]))[0] == 0.0
    sp = _eval(O, 1, np.float32([0.0, -1.0, np.inf, np.nan]))
    assert sp[0] == -np.inf and np.isnan(sp[1]) and sp[2] == np.inf and np.isnan(sp[3])


def test_log2_exp2_pow(O):
    x = np.exp(np.linspace(-20, 20, 50001)).astype(np.float32)
    assert _ulps(_eval(O, 2, x), np.log2(x.astype(np.float64))).max() <= 4 or \
        np.abs(_eval(O, 2, x) - np.log2(x.astype(np.float64))).max() < 1e-6
    y = np.linspace(-120, 120, 50001).astype(np.float32)
    assert _ulps(_eval(O, 3, y), np.exp2(y.astype(np.float64))).max() <= 3.0
    # exact powers of two
    k = np.arange(-126, 128, dtype=np.float32)
    assert np.array_equal(_eval(O, 3, k), np.ldexp(np.float64(1.0), k.astype(np.int64)).astype(np.float32))
    a = np.exp(np.linspace(-5, 9, 2001)).astype(np.float32)
    b = np.linspace(0, 1, 2001).astype(np.float32)
    got = _eval(O, 4, a, b)
    ref = np.power(a.astype(np.float64), b.astype(np.float64))
    assert (np.abs(got - ref) / ref).max() < 2e-6
    # pow(x, 0) == 1 and pow(x, 1) == x exactly (confidence-weighted update with fraction 0/1)
    assert np.array_equal(_eval(O, 4, a, np.zeros_like(a)), np.ones_like(a))
    assert np.array_equal(_eval(O, 4, a, np.ones_like(a)), a)


def test_half_round_trip_matches_numpy(O):
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.normal(0, 1, 100000) * 10.0 ** rng.integers(-9, 6, 100000),
                        [65504.0, 65519.99, 65520.0, 1e9, 6.1e-5, 5.96e-8, 2.98e-8, 2.99e-8, 0.0, -0.0, 0.1]]
                       ).astype(np.float32)
    got = _eval(O, 5, x)
    with np.errstate(over="ignore"):
        ref = x.astype(np.float16).astype(np.float32)
    same = (got == ref) | (np.isnan(got) & np.isnan(ref))
    assert same.all(), x[~same][:5]


def test_round_and_floor_conversions(O):
    x = np.float32([0.5, 1.5, 2.5, -0.5, -1.5, 32767.4, 32768.0, -32769.0, 70000.0, 3e9, np.nan])
    r = _eval(O, 6, x)
    # RNE, then int32 -> int16 wrap; out of int32 range or NaN -> 0x80000000 -> 0
    assert list(r) == [0, 2, 2, 0, -2, 32767, -32768, 32767, 4464, 0, 0]
    f = _eval(O, 7, np.float32([-0.5, 0.99999994, 299.99997, -1e10]))
    assert list(f) == [-1, 0, 299, -2147483648.0]


def test_table_log_and_reciprocal_accuracy(O):
    """hw_log_rcp_tab (the EM's log and reciprocal of bin probabilities, positive
    normal floats): within 2 ulp of the float64 log where |log x| >= 0.5, within 1e-7
    absolute below that, and within 2 ulp of 1/x, over every binade and every table
    interval (x = 2^e (1 + j / 2^23) at a stride of the fraction)."""
    lo = np.float32(1.1754944e-38).view(np.uint32)
    hi = np.float32(1.7e38).view(np.uint32)
    bits = np.concatenate([np.arange(lo, hi, 7919, dtype=np.uint32),
                           np.uint32([lo, hi, 0x3f800000, 0x3f7fffff, 0x3f800001, 0x3fffffff])])
    x = bits.view(np.float32)
    lg = _eval(O, 9, x).astype(np.float64)
    rcp = _eval(O, 10, x)
    ref = np.log(x.astype(np.float64))
    big = np.abs(ref) >= 0.5
    assert _ulps(lg[big].astype(np.float32), ref[big]).max() <= 2
    assert np.abs(lg[~big] - ref[~big]).max() <= 1e-7
    assert _ulps(rcp, 1.0 / x.astype(np.float64)).max() <= 2


# ---------------------------------------------------------------------------
# The reference's own transcendental routines (Cephes exp / log of its SSE host
# build, common/simd/sse_mathfun.h via m512f_SSE.h:444-466 and xcompile.h:148-151)
# and the correctly rounded exp2 / log2 / pow standing in for its libm calls.
# ---------------------------------------------------------------------------

def test_cephes_exp_accuracy_and_sse_limits(O):
    """Cephes expf is within one ulp of exp over its range (a wrong coefficient,
    a dropped C2 reduction term or a misplaced floor correction all exceed it);
    outside the clamp at +-88.3762626647949 the SSE routine saturates: 0 below,
    a finite value above, and NaN (which the SSE min/max rule turns into the
    upper clamp) gives the same value as the clamp."""
    x = np.concatenate([np.linspace(-87, 88, 400001, dtype=np.float32),
                        np.float32([0.0, -0.0, 1e-30, -1e-30, 0.5, -0.5, 1.0, -1.0])])
    got = _eval(O, 11, x)
    assert _ulps(got, np.exp(x.astype(np.float64))).max() <= 1.0
    assert _eval(O, 11, np.float32([0.0]))[0] == 1.0
    hi = np.float32(88.3762626647949)
    sp = _eval(O, 11, np.float32([100.0, hi, np.nan, -100.0, -88.3762626647949, np.inf, -np.inf]))
    assert sp[0] == sp[1] == sp[2] and np.isfinite(sp[0]) and sp[0] > 1e38
    assert sp[3] == 0.0 and sp[4] == 0.0 and sp[6] == 0.0 and sp[5] == sp[0]
    # just above the lower clamp the result is a subnormal 2^-126 multiple, not 0
    assert 0.0 < _eval(O, 11, np.float32([-87.5]))[0] < np.float32(1.2e-38)


def test_cephes_log_accuracy_and_sse_limits(O):
    """Cephes logf is within one ulp of log where |log x| >= 0.01 and on every
    float within 4000 ulp of 1 (the r = 2m - 1 branch below sqrt(1/2) included);
    x <= 0 gives NaN, 0 gives -inf, subnormals and NaN are raised to FLT_MIN
    first (the SSE max rule), and +inf returns log of the largest binade's
    mantissa bits, as the bit manipulation does."""
    x = np.exp(np.linspace(-87, 88, 400001)).astype(np.float32)
    got = _eval(O, 12, x).astype(np.float64)
    ref = np.log(x.astype(np.float64))
    big = np.abs(ref) >= 0.01
    assert _ulps(got[big].astype(np.float32), ref[big]).max() <= 1.0
    near1 = np.float32(1) + np.arange(-4000, 4000, dtype=np.float32) * np.float32(2 ** -23)
    nz = near1 != 1
    assert _ulps(_eval(O, 12, near1)[nz], np.log(near1[nz].astype(np.float64))).max() <= 1.0
    assert _eval(O, 12, np.float32([1.0]))[0] == 0.0
    sp = _eval(O, 12, np.float32([0.0, -0.0, -1.0, np.nan, 1e-40, 1.1754944e-38, np.inf]))
    assert sp[0] == -np.inf and sp[1] == -np.inf and np.isnan(sp[2])
    assert sp[3] == sp[4] == sp[5] and abs(sp[5] - np.log(1.1754944e-38)) < 1e-5
    assert np.isfinite(sp[6]) and abs(sp[6] - 88.7228) < 1e-3
    # both sides of the sqrt(1/2) branch point agree with log to an ulp
    m = np.float32([0.70710677, 0.7071067, 0.70710683, 1.4142135, 1.4142137])
    assert _ulps(_eval(O, 12, m), np.log(m.astype(np.float64))).max() <= 1.0


def test_double_rounded_exp2_log2_pow_are_correctly_rounded(O):
    """hw_dexp2f / hw_dlog2f / hw_dpowf (the kernels' stand-ins for the libm
    exp2f / log2f / powf the reference calls) equal the float64 result rounded
    once, and agree with glibc's float functions except on a small fraction of
    arguments where those are off by exactly one ulp."""
    y = np.linspace(-120, 120, 200001).astype(np.float32)
    x = np.exp(np.linspace(-80, 80, 200001)).astype(np.float32)
    a = np.repeat(np.exp(np.linspace(-5, 11, 401)), 401).astype(np.float32)
    b = np.tile(np.linspace(0, 1, 401), 401).astype(np.float32)
    cases = [(16, 13, y, None, np.exp2(y.astype(np.float64))),
             (17, 14, x, None, np.log2(x.astype(np.float64))),
             (18, 15, a, b, np.power(a.astype(np.float64), b.astype(np.float64)))]
    for mine, libm, u, v, ref in cases:
        got = _eval(O, mine, u, v)
        assert np.array_equal(got, ref.astype(np.float32)), mine
        lm = _eval(O, libm, u, v)
        diff = got != lm
        assert diff.mean() < 2e-3, (mine, diff.mean())
        assert _ulps(lm[diff], got[diff].astype(np.float64)).max() <= 1.0
