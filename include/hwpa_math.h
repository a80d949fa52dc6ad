/*
 * hwpa_math.h -- deterministic single-precision math shared by the gfx950
 * kernels and their host-side checkers.
 *
 * Every routine here is written with IEEE-754 basic operations only
 * (add, sub, mul, explicit fused multiply-add, correctly rounded div and
 * sqrt, floor, rint, ldexp, bit casts), so a translation unit built with
 * -ffp-contract=off produces the same bits on the host and on the GPU
 * (an explicit fmaf is a single correctly rounded operation on both).  The transcendental functions are derived
 * from their Taylor / atanh series, not from any vendor library:
 *   exp(x)  = 2^k * exp(r),  r = x - k ln2 (two-part ln2),  |r| <= ln2/2,
 *             exp(r) by its Taylor series to r^7 (truncation < 6e-9 rel),
 *             2^k applied with an exact ldexp.
 *   log(x)  = e ln2 + 2 atanh(s),  s = (m-1)/(m+1),  m in [sqrt(1/2), sqrt(2)),
 *             atanh series to s^13 (truncation < 2e-10).
 * Accuracy is a few ulp; tests/test_math.py bounds the error against libm.
 * Where the reference's own transcendental bits are cheap to reproduce, the
 * kernels use its routines instead: hw_cephes_expf / hw_cephes_logf restate
 * the Cephes exp / log of its SSE host build, and hw_dexp2f / hw_dlog2f /
 * hw_dpowf give the libm values it calls for exp2, log2 and pow.
 *
 * Conversions follow the reference's SSE LaneArray semantics:
 *   float -> int32 : round-to-nearest-even, 0x80000000 when out of range or NaN
 *                    (_mm_cvtps_epi32 after _mm_round_ps, common/simd/m512f_SSE.h:126-132)
 *   int32 -> int16 : wrap (common/simd/m512s_SSE.h:101-113)
 *   float -> half  : IEEE round-to-nearest-even (PBHalf = cuda half)
 */
#ifndef HWPA_MATH_H
#define HWPA_MATH_H

#include <stdint.h>
#if !defined(__HIPCC__)
#include <math.h>
#endif

#if defined(__HIPCC__)
#define HWPA_HD __host__ __device__ inline
#else
#define HWPA_HD inline
#endif

HWPA_HD uint32_t hw_f2u(float f) { return __builtin_bit_cast(uint32_t, f); }
HWPA_HD float hw_u2f(uint32_t u) { return __builtin_bit_cast(float, u); }

HWPA_HD float hw_floorf(float x) { return __builtin_floorf(x); }
HWPA_HD float hw_rintf(float x) { return __builtin_rintf(x); }
HWPA_HD float hw_sqrtf(float x) { return __builtin_sqrtf(x); }
HWPA_HD float hw_fmaf(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
HWPA_HD float hw_fabsf(float x) { return hw_u2f(hw_f2u(x) & 0x7fffffffu); }
HWPA_HD bool hw_isnan(float x) { return (hw_f2u(x) & 0x7fffffffu) > 0x7f800000u; }
HWPA_HD bool hw_isinf(float x) { return (hw_f2u(x) & 0x7fffffffu) == 0x7f800000u; }
HWPA_HD bool hw_isfinite(float x) { return (hw_f2u(x) & 0x7f800000u) != 0x7f800000u; }

/* SSE min/max: the second operand is returned when either is NaN
 * (_mm_min_ps / _mm_max_ps). */
HWPA_HD float hw_minf(float a, float b) { return (a < b) ? a : b; }
HWPA_HD float hw_maxf(float a, float b) { return (a > b) ? a : b; }
HWPA_HD float hw_clampf(float v, float lo, float hi) { return hw_minf(hw_maxf(v, lo), hi); }
/* hw_maxf(a, b) for b positive and not NaN, in one instruction on the GPU: the IEEE
 * maximum returns b for a NaN a, as the comparison does, and the two operands are never
 * zeros of opposite sign. */
HWPA_HD float hw_max_pos(float a, float b) { return __builtin_fmaxf(a, b); }
/* satlin (common/NumericUtil.h:70-76). */
HWPA_HD float hw_satlin(float a, float b, float x)
{
    const float r = (x - a) / (b - a);
    return hw_minf(hw_maxf(r, 0.0f), 1.0f);
}

/* float -> int32 with the x86 cvtps semantics for out-of-range values. */
HWPA_HD int32_t hw_cvt_i32(float x)
{
    if (!(x >= -2147483648.0f && x < 2147483648.0f)) return (int32_t)0x80000000u;
    return (int32_t)x;
}
HWPA_HD int32_t hw_round_i32(float x) { return hw_cvt_i32(hw_rintf(x)); }
HWPA_HD int32_t hw_floor_i32(float x) { return hw_cvt_i32(hw_floorf(x)); }
HWPA_HD int16_t hw_wrap16(int32_t v) { return (int16_t)(uint16_t)(uint32_t)v; }
HWPA_HD int16_t hw_round_i16(float x) { return hw_wrap16(hw_round_i32(x)); }

/* 2^k for integer k in [-149, 127]. */
HWPA_HD float hw_pow2i(int k)
{
    if (k >= -126) return hw_u2f((uint32_t)(k + 127) << 23);
    return hw_u2f(1u << (k + 149));
}

/* x * 2^k, exact (one rounding for subnormal results).  The GPU has it as one
 * instruction; the host builds the bits when the result stays normal. */
HWPA_HD float hw_ldexpf(float x, int k)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_ldexpf(x, k);
#else
    const uint32_t u = hw_f2u(x);
    const int e = (int)((u >> 23) & 0xffu);
    if (e != 0 && e != 255 && k > -e && k < 255 - e) return hw_u2f(u + ((uint32_t)k << 23));
    return __builtin_ldexpf(x, k);
#endif
}

/* Clamp for reduction arguments; a NaN argument gives an unspecified finite
 * value, so callers restore NaN themselves. */
HWPA_HD float hw_clamp_arg(float x, float lo, float hi)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_fminf(__builtin_fmaxf(x, lo), hi);
#else
    return hw_minf(hw_maxf(x, lo), hi);
#endif
}

/* exp(r) for |r| <= ln2/2: Taylor polynomial to r^7 in Horner form. */
HWPA_HD float hw_exp_poly(float r)
{
    float p = 1.0f / 5040.0f;
    p = hw_fmaf(p, r, 1.0f / 720.0f);
    p = hw_fmaf(p, r, 1.0f / 120.0f);
    p = hw_fmaf(p, r, 1.0f / 24.0f);
    p = hw_fmaf(p, r, 1.0f / 6.0f);
    p = hw_fmaf(p, r, 0.5f);
    p = hw_fmaf(p, r, 1.0f);
    return hw_fmaf(p, r, 1.0f);
}

/* exp(x): k = rint(x / ln2), r = x - k ln2 (Cody-Waite, |r| <= ln2/2),
 * exp(r) by its Taylor polynomial, scaled by 2^k with an exact ldexp (which
 * also produces the correctly rounded subnormal / overflow results).  The
 * argument is clamped to +-200 so k stays in int range; NaN propagates. */
HWPA_HD float hw_expf(float x)
{
    const float xc = hw_clamp_arg(x, -200.0f, 200.0f);
    const float kf = hw_rintf(xc * 1.44269502162933349609375f);
    /* ln2 = ln2hi + ln2lo; with fused steps the reduction keeps ~48 bits. */
    const float ln2hi = 0.693145751953125f;
    const float ln2lo = 1.428606765330187045e-06f;
    const float r = hw_fmaf(-kf, ln2lo, hw_fmaf(-kf, ln2hi, xc));
    const float v = hw_ldexpf(hw_exp_poly(r), (int)kf);
    return hw_isnan(x) ? x : v;
}

/* log(m) for m in [sqrt(1/2), sqrt(2)) and the binary exponent e of the
 * positive normal float with bits u, offset by e0. */
HWPA_HD float hw_log_reduce_bits(uint32_t u, int e0, int* e_out)
{
    int e = e0 + (int)((u >> 23) & 0xffu) - 127;
    float m = hw_u2f((u & 0x007fffffu) | 0x3f800000u);
    if (m > 1.41421353816986083984375f) {
        m = m * 0.5f;
        e += 1;
    }
    const float s = (m - 1.0f) / (m + 1.0f);
    const float s2 = s * s;
    float q = 1.0f / 13.0f;
    q = hw_fmaf(q, s2, 1.0f / 11.0f);
    q = hw_fmaf(q, s2, 1.0f / 9.0f);
    q = hw_fmaf(q, s2, 1.0f / 7.0f);
    q = hw_fmaf(q, s2, 1.0f / 5.0f);
    q = hw_fmaf(q, s2, 1.0f / 3.0f);
    q = q * s2;
    *e_out = e;
    return hw_fmaf(s + s, q, s + s);
}

/* Returns log(m) for m in [sqrt(1/2), sqrt(2)) and the binary exponent e. */
HWPA_HD float hw_log_reduce(float x, int* e_out)
{
    uint32_t u = hw_f2u(x);
    int e0 = 0;
    if (u < 0x00800000u) {           /* subnormal: scale by 2^24 */
        x = x * 16777216.0f;
        u = hw_f2u(x);
        e0 = -24;
    }
    return hw_log_reduce_bits(u, e0, e_out);
}

HWPA_HD float hw_logf(float x)
{
    if (hw_isnan(x)) return x;
    if (x < 0.0f) return hw_u2f(0x7fc00000u);
    if (x == 0.0f) return hw_u2f(0xff800000u);
    if (hw_isinf(x)) return x;
    int e;
    const float lm = hw_log_reduce(x, &e);
    const float ef = (float)e;
    const float ln2hi = 0.693145751953125f;
    const float ln2lo = 1.428606765330187045e-06f;
    return hw_fmaf(ef, ln2hi, hw_fmaf(ef, ln2lo, lm));
}

/* Table-driven log and reciprocal of a positive normal finite x (the EM's bin
 * probabilities).  x = 2^e m with m in [1, 2); k = the top 6 bits of m's fraction,
 * c_k = 1 + (k + 1/2) / 64, rc_k = 1 / c_k rounded, lc_k = -hw_logf(rc_k) (the
 * table, hw_logtab_entry); r = m rc_k - 1 in one fused step, |r| < 2^-7 + 2^-24:
 *   log x = e ln2 + lc_k + log1p(r),  log1p(r) = r - r^2/2 + r^3/3 - r^4/4
 *           (truncation below |r|^5 / 5 < 1e-11),
 *   1 / x = 2^-e rc_k (1 - r + r^2 - r^3 + r^4)   (truncation below |r|^5).
 * A few ulp, no division. */
#define HW_LOGTAB_N 64
HWPA_HD void hw_logtab_entry(int k, float* rc, float* lc)
{
    const float c = 1.0f + ((float)k + 0.5f) / 64.0f;
    *rc = 1.0f / c;
    *lc = -hw_logf(*rc);
}

HWPA_HD void hw_log_rcp_tab(float x, const float* rcTab, const float* lcTab, float* logOut, float* rcpOut)
{
    const uint32_t u = hw_f2u(x);
    const int e = (int)((u >> 23) & 0xffu) - 127;
    const uint32_t k = (u >> 17) & (uint32_t)(HW_LOGTAB_N - 1);
    const float m = hw_u2f((u & 0x007fffffu) | 0x3f800000u);
    const float rc = rcTab[k];
    const float r = hw_fmaf(m, rc, -1.0f);
    float p = hw_fmaf(-0.25f, r, 1.0f / 3.0f);
    p = hw_fmaf(p, r, -0.5f);
    p = hw_fmaf(p, r, 1.0f);
    const float ef = (float)e;
    const float ln2hi = 0.693145751953125f;
    const float ln2lo = 1.428606765330187045e-06f;
    *logOut = hw_fmaf(ef, ln2hi, hw_fmaf(ef, ln2lo, lcTab[k] + p * r));
    float q = r - 1.0f;
    q = hw_fmaf(q, r, 1.0f);
    q = hw_fmaf(q, r, -1.0f);
    q = hw_fmaf(q, r, 1.0f);
    *rcpOut = hw_ldexpf(rc * q, -e);
}


HWPA_HD float hw_log2f(float x)
{
    if (hw_isnan(x)) return x;
    if (x < 0.0f) return hw_u2f(0x7fc00000u);
    if (x == 0.0f) return hw_u2f(0xff800000u);
    if (hw_isinf(x)) return x;
    int e;
    const float lm = hw_log_reduce(x, &e);
    return hw_fmaf(lm, 1.44269502162933349609375f, (float)e);
}

HWPA_HD float hw_exp2f(float x)
{
    const float xc = hw_clamp_arg(x, -300.0f, 300.0f);
    const float kf = hw_rintf(xc);
    const float r = (xc - kf) * 0.693147182464599609375f;   /* xc - kf is exact */
    const float v = hw_ldexpf(hw_exp_poly(r), (int)kf);
    return hw_isnan(x) ? x : v;
}

/* The reference's LaneArray exp and log as its x86 host build evaluates them:
 * m512f exp/log (common/simd/m512f_SSE.h:444-466) call _mm_exp_ps/_mm_log_ps
 * (xcompile.h:148-151), i.e. sse_mathfun's exp_ps/log_ps, the Cephes library's
 * single-precision expf/logf (S. L. Moshier) on SSE registers.  Restated here
 * per element with the same basic operations in the same order, no fused steps,
 * so host and GPU reproduce its bits:
 *   exp(x): x clamped to +-88.3762626647949, n = floor(x log2(e) + 1/2) by
 *           truncation and correction, r = (x - n C1) - n C2 with
 *           C1 + C2 = ln 2, p(r) = 1 + r + r^2 (degree-5 minimax polynomial),
 *           times 2^n built from its exponent bits.
 *   log(x): x raised to FLT_MIN, x = m 2^e with m in [1/2, 1); if m < sqrt(1/2)
 *           then r = 2m - 1, e -= 1, else r = m - 1; log = r + r^3 P(r)
 *           - r^2/2 + e (q1 + q2) with q1 + q2 = ln 2; x <= 0 gives NaN
 *           (all bits set), x == 0 gives -inf.
 * A NaN argument of exp returns exp(88.376...) (+inf), of log returns
 * log(FLT_MIN), as the SSE min/max rule makes them.  Coefficients are the
 * Cephes ones, rounded from double as the reference's constant tables are. */
HWPA_HD float hw_cephes_expf(float x)
{
    x = hw_minf(x, (float)88.3762626647949);
    x = hw_maxf(x, (float)-88.3762626647949);
    float fx = x * (float)1.44269504088896341;
    fx = fx + 0.5f;
    float t = (float)(int32_t)fx;
    if (t > fx) t = t - 1.0f;
    fx = t;
    float p = fx * (float)0.693359375;
    const float q = fx * (float)-2.12194440e-4;
    x = x - p;
    x = x - q;
    const float z = x * x;
    float y = (float)1.9875691500E-4;
    y = y * x;
    y = y + (float)1.3981999507E-3;
    y = y * x;
    y = y + (float)8.3334519073E-3;
    y = y * x;
    y = y + (float)4.1665795894E-2;
    y = y * x;
    y = y + (float)1.6666665459E-1;
    y = y * x;
    y = y + (float)5.0000001201E-1;
    y = y * z;
    y = y + x;
    y = y + 1.0f;
    p = hw_u2f((uint32_t)((int32_t)fx + 127) << 23);
    return y * p;
}

HWPA_HD float hw_cephes_logf(float x)
{
    const bool invalid = x <= 0.0f;
    const bool zero = x == 0.0f;
    x = hw_maxf(x, hw_u2f(0x00800000u));
    const uint32_t u = hw_f2u(x);
    float e = (float)((int32_t)(u >> 23) - 127);
    x = hw_u2f((u & ~0x7f800000u) | 0x3f000000u);
    e = e + 1.0f;
    const bool small = x < (float)0.707106781186547524;
    const float t = small ? x : 0.0f;
    x = x - 1.0f;
    e = e - (small ? 1.0f : 0.0f);
    x = x + t;
    const float z = x * x;
    float y = (float)7.0376836292E-2;
    y = y * x;
    y = y + (float)-1.1514610310E-1;
    y = y * x;
    y = y + (float)1.1676998740E-1;
    y = y * x;
    y = y + (float)-1.2420140846E-1;
    y = y * x;
    y = y + (float)1.4249322787E-1;
    y = y * x;
    y = y + (float)-1.6668057665E-1;
    y = y * x;
    y = y + (float)2.0000714765E-1;
    y = y * x;
    y = y + (float)-2.4999993993E-1;
    y = y * x;
    y = y + (float)3.3333331174E-1;
    y = y * x;
    y = y * z;
    float s = e * (float)-2.12194440e-4;
    y = y + s;
    s = z * 0.5f;
    y = y - s;
    s = e * (float)0.693359375;
    x = x + y;
    x = x + s;
    if (invalid) x = hw_u2f(0xffffffffu);
    if (zero) x = hw_u2f(0xff800000u);
    return x;
}

/* The reference calls libm per element for LaneArray exp2, log2 and pow
 * (common/simd/xcompile.h:180-197, m512f.h:34-41).  These evaluate the
 * double-precision function and round once to float: the correctly rounded
 * float result (the host and the GPU double functions are both within an ulp
 * of the exact value, far inside the float rounding step), which is libm's
 * float result except on the under 0.1% of arguments where glibc's exp2f /
 * log2f / powf are off by one ulp. */
HWPA_HD float hw_dexp2f(float x) { return (float)exp2((double)x); }
HWPA_HD float hw_dlog2f(float x) { return (float)log2((double)x); }
HWPA_HD float hw_dpowf(float x, float y) { return (float)pow((double)x, (double)y); }

/* Logit of a probability threshold p in [0, 1), evaluated on the host in double
 * (-inf for p = 0, so that every finite network output passes it). */
static inline float hw_gate_logit(float p)
{
    return p > 0.0f ? (float)log((double)p / (1.0 - (double)p)) : -__builtin_inff();
}

/* x^y for x > 0; exact for y == 0 and y == 1. */
HWPA_HD float hw_powf(float x, float y)
{
    if (y == 0.0f) return 1.0f;
    if (y == 1.0f) return x;
    if (hw_isnan(x) || hw_isnan(y)) return hw_u2f(0x7fc00000u);
    if (x == 1.0f) return 1.0f;
    if (x < 0.0f) return hw_u2f(0x7fc00000u);
    if (x == 0.0f) return (y > 0.0f) ? 0.0f : hw_u2f(0x7f800000u);
    return hw_exp2f(y * hw_log2f(x));
}

/* IEEE binary16 <-> binary32, round to nearest even. */
HWPA_HD uint16_t hw_f2h(float f)
{
    const uint32_t u = hw_f2u(f);
    const uint32_t sign = (u >> 16) & 0x8000u;
    const uint32_t a = u & 0x7fffffffu;
    if (a > 0x7f800000u) return (uint16_t)(sign | 0x7e00u | ((a >> 13) & 0x3ffu));
    if (a >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);   /* rounds to >= 65520 */
    if (a >= 0x38800000u) {                                     /* normal half */
        const uint32_t mant = a & 0x7fffffu;
        const uint32_t exp = (a >> 23) - 112u;
        uint32_t h = (exp << 10) | (mant >> 13);
        const uint32_t rem = mant & 0x1fffu;
        if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h += 1u;
        return (uint16_t)(sign | h);
    }
    if (a < 0x33000000u) return (uint16_t)sign;                 /* < 2^-25 rounds to 0 */
    /* subnormal half: value = mant * 2^-24 */
    const uint32_t e = a >> 23;                                  /* 102..112 */
    const uint32_t mant = (a & 0x7fffffu) | 0x800000u;
    const uint32_t shift = 126u - e;                             /* 14..24 */
    uint32_t h = mant >> shift;
    const uint32_t rem = mant & ((1u << shift) - 1u);
    const uint32_t half = 1u << (shift - 1u);
    if (rem > half || (rem == half && (h & 1u))) h += 1u;
    return (uint16_t)(sign | h);
}

HWPA_HD float hw_h2f(uint16_t h)
{
    const uint32_t sign = ((uint32_t)h & 0x8000u) << 16;
    const uint32_t exp = ((uint32_t)h >> 10) & 0x1fu;
    const uint32_t mant = (uint32_t)h & 0x3ffu;
    if (exp == 0x1fu) return hw_u2f(sign | 0x7f800000u | (mant << 13));
    if (exp != 0u) return hw_u2f(sign | ((exp + 112u) << 23) | (mant << 13));
    if (mant == 0u) return hw_u2f(sign);
    /* subnormal: mant * 2^-24 is exact in float */
    const float v = (float)mant * 5.9604644775390625e-08f;
    return sign ? -v : v;
}

/* bfloat16 <-> binary32: round to nearest even on the top 16 bits, NaN kept
 * quiet.  Used for the initialiser network's weights and activations. */
HWPA_HD uint16_t hw_f2bf(float f)
{
    const uint32_t u = hw_f2u(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
/* hw_f2bf for a value known not to be NaN (a ReLU output, (z > 0 ? z : 0), never is):
 * the same round-to-nearest-even without the NaN case. */
HWPA_HD uint16_t hw_f2bf_nonnan(float f)
{
    const uint32_t u = hw_f2u(f);
    return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
HWPA_HD float hw_bf2f(uint16_t b) { return hw_u2f((uint32_t)b << 16); }

/* Round a float through binary16 and back. */
HWPA_HD float hw_half_round(float f) { return hw_h2f(hw_f2h(f)); }

/* Histogram bin without a divide (TraceHistogramAccumHost::AddBlock's
 * floorCastInt((x - lo) / binSize), dataTypes/UHistogramSimd.h:199-229, for the
 * samples where it can be decided exactly).  The bin is floor(Q), Q = RN(a / bs),
 * a = RN(x - lo).  With q = RN(a * RN(1 / bs)), within 2^-23 |a / bs| of a / bs,
 * |q| < 2048, k = rint(q) and r = fma(-k, bs, a): for bs > 0, r has the sign of
 * a - k bs exactly (the exact value is a multiple of the smaller of ulp(a) and
 * ulp(bs), so it cannot round to 0 unless it is 0), and
 *   r >= 0: a / bs >= k, so Q >= k, and a / bs < k + 0.501, so Q < k + 1: bin k;
 *   r < 0:  a / bs < k, so Q <= k, and Q > k - 1: bin k - 1, unless Q rounded up
 *           to k (a / bs within half an ulp of k below it, at most |k| 2^-24) or,
 *           for k = 0, to -0 (|a / bs| < 2^-149).
 * Both exceptions lie inside -r <= bs |k| 2^-22 + bs 2^-100.  hw_shortbin returns
 * false for them, for bs <= 0 or non-finite (a NaN reciprocal fails |q| < 2048),
 * for |q| >= 2048 or NaN and for bins outside [0, nbins): those samples need the
 * divide.  The conditions combine bitwise, so the test has no branches. */
typedef struct hw_shortbin_z {
    float lo, bs, inv, bsc, bst;   /* inv = 1 / bs or NaN, bsc = bs 2^-22, bst = bs 2^-100 */
} hw_shortbin_z;

HWPA_HD hw_shortbin_z hw_shortbin_make(float lo, float bs)
{
    hw_shortbin_z zb;
    zb.lo = lo;
    zb.bs = bs;
    zb.inv = (bs > 0.0f && bs < hw_u2f(0x7f800000u)) ? 1.0f / bs : hw_u2f(0x7fc00000u);
    zb.bsc = bs * 2.384185791015625e-07f;
    zb.bst = bs * 7.888609052210118e-31f;
    return zb;
}

HWPA_HD bool hw_shortbin(int32_t x, const hw_shortbin_z& zb, int32_t nbins, int32_t* bin)
{
    const float a = (float)x - zb.lo;
    const float q = a * zb.inv;
    const float k = hw_rintf(q);
    const float r = hw_fmaf(-k, zb.bs, a);
    const float band = hw_fmaf(hw_fabsf(k), zb.bsc, zb.bst);
    const bool inq = hw_fabsf(q) < 2048.0f;
    const bool below = r < 0.0f;
    const int32_t b = (int32_t)(inq ? k : 0.0f) - (below ? 1 : 0);
    *bin = b;
    const bool amb = below & (r >= -band);
    return inq & !amb & ((uint32_t)b < (uint32_t)nbins);
}

/* x / d for a constant divisor d with rd = RN(1 / d), without a divide: the
 * product q0 = RN(x rd), its exact residual x - q0 d (one fused multiply-add)
 * and one correction step q0 + RN(residual) rd (Markstein's theorem for a
 * correctly rounded reciprocal).  For d = RN(sqrt 2) and RN(sqrt 3) (the ROI
 * filter's divisors) the result equals the IEEE quotient for every finite x with
 * |x| >= 2^-100, for every zero up to its sign, and for infinities and NaN (the
 * final select); below 2^-100 the residual is no longer exact.  The ROI sums it
 * serves are 0 or at least 2^-31 in magnitude (int16 samples times 1/sigma with
 * sigma^2 a finite half). */
HWPA_HD float hw_div_by_const(float x, float d, float rd)
{
    const float q0 = x * rd;
    const float e = hw_fmaf(-q0, d, x);
    const float q1 = hw_fmaf(e, rd, q0);
    return (hw_fabsf(q0) < hw_u2f(0x7f800000u)) ? q1 : q0;
}

/* The ROI backward recursion over one word of 16 two-bit fields (field k in
 * bits 2k, 2k+1: the forward pass's roi and mid bits), from field 15 down to
 * field 0: final_k = roi_k | (mid_k & final_{k+1}), with final_16 = *carry.
 * Bit-reversed, the recurrence runs towards higher bits and is the carry chain
 * of one addition: roi generates, mid propagates, and the bit between two
 * fields passes the carry on.  Returns final_k in bit 2k (odd bits zero);
 * *carry becomes final_0.  Fields holding 0 give 0 and pass on no carry. */
HWPA_HD uint32_t hw_bitrev32(uint32_t v)
{
#if defined(__clang__)
    return __builtin_bitreverse32(v);
#else
    v = ((v >> 1) & 0x55555555u) | ((v & 0x55555555u) << 1);
    v = ((v >> 2) & 0x33333333u) | ((v & 0x33333333u) << 2);
    v = ((v >> 4) & 0x0F0F0F0Fu) | ((v & 0x0F0F0F0Fu) << 4);
    v = ((v >> 8) & 0x00FF00FFu) | ((v & 0x00FF00FFu) << 8);
    return (v >> 16) | (v << 16);
#endif
}

HWPA_HD uint32_t hw_roi_backward_word(uint32_t w, uint32_t* carry)
{
    const uint32_t rv = hw_bitrev32(w);               /* roi_k at 31-2k, mid_k at 30-2k */
    const uint32_t g = rv & 0xAAAAAAAAu;
    const uint32_t p = (rv << 1) & 0xAAAAAAAAu;
    const uint32_t a = g | p | 0x55555555u;
    const uint32_t s = a + g + *carry;                /* the carry enters at bit 0 */
    const uint32_t cin = s ^ a ^ g;                   /* carry into every bit */
    const uint32_t out = hw_bitrev32(g | (p & cin));
    *carry = out & 1u;
    return out;
}

#endif /* HWPA_MATH_H */
