/*
 * hwpa_tracegen.h -- seeded synthetic int16 trace generator.
 *
 * Each ZMW is an independent stream: every random draw is a hash of
 * (seed, zmw, frame, stream id, attempt), so the host and the GPU produce
 * identical traces and any chunk can be generated from the carried
 * per-ZMW state alone.
 *
 * Model per frame t of ZMW z:
 *   value = mu + slope * t / drift_chunk + step(t) + sigma * N(0,1)
 *           + [in pulse] amp[a] * (1 + cv * N(0,1))
 * rounded to nearest-even and saturated to int16.  Optionally all amplitudes
 * of a ZMW are scaled by a per-ZMW factor drawn from a signal-to-noise range.  The pulse process is a
 * two-state renewal chain with per-frame exits, so gaps and widths are
 * geometric with means mean_ipd and mean_pulse_width; the analog a is drawn
 * uniformly at each pulse start.  Normals use Marsaglia's polar method.
 */
#ifndef HWPA_TRACEGEN_H
#define HWPA_TRACEGEN_H

#include "hwpa_math.h"
#include "hwpa_types.h"

HWPA_HD uint64_t hw_mix64(uint64_t x)
{
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ull;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebull;
    x ^= x >> 31;
    return x;
}

HWPA_HD uint64_t hw_hash(uint64_t seed, uint64_t zmw, uint64_t ctr, uint32_t stream)
{
    uint64_t h = hw_mix64(seed ^ 0x9e3779b97f4a7c15ull);
    h = hw_mix64(h ^ (zmw * 0xd1b54a32d192ed03ull));
    h = hw_mix64(h ^ ((uint64_t)stream * 0xa0761d6478bd642full));
    h = hw_mix64(h ^ (ctr * 0x8cb92ba72f3d8dd7ull));
    return h;
}

/* Uniform in (0, 1) with 24 random bits. */
HWPA_HD float hw_uniform(uint64_t h)
{
    return ((float)(uint32_t)(h >> 40) + 0.5f) * 5.9604644775390625e-08f;
}

HWPA_HD float hw_normal(uint64_t seed, uint64_t zmw, uint64_t ctr, uint32_t stream)
{
    for (uint32_t attempt = 0; attempt < 32u; ++attempt) {
        const uint64_t h = hw_hash(seed, zmw, ctr, stream + (attempt << 8));
        const float u = 2.0f * ((float)(uint32_t)(h >> 40) + 0.5f) * 5.9604644775390625e-08f - 1.0f;
        const float v = 2.0f * ((float)(uint32_t)((h >> 16) & 0xffffffu) + 0.5f) * 5.9604644775390625e-08f - 1.0f;
        const float s = u * u + v * v;
        if (s < 1.0f && s > 0.0f) {
            return u * hw_sqrtf((-2.0f * hw_logf(s)) / s);
        }
    }
    return 0.0f;
}

enum { HW_STREAM_NOISE = 1, HW_STREAM_EXIT = 2, HW_STREAM_ANALOG = 3, HW_STREAM_AMP = 4,
       HW_STREAM_SLOPE = 5, HW_STREAM_STEP = 6, HW_STREAM_EMPTY = 7, HW_STREAM_SNR = 8 };

/* Amplitude scale of ZMW z: 1, or snr * sigma / amplitude[0] with
 * snr ~ U[snr_lo, snr_hi] when snr_hi > 0. */
HWPA_HD float hw_tracegen_amp_scale(const hwpa_tracegen_params* p, uint64_t z)
{
    if (!(p->snr_hi > 0.0f)) return 1.0f;
    const float u = hw_uniform(hw_hash(p->seed, z, 0, HW_STREAM_SNR));
    const float snr = p->snr_lo + (p->snr_hi - p->snr_lo) * u;
    return (snr * p->baseline_sigma) / p->amplitude[0];
}

/* Initialise the carried state of ZMW z (slot i of its lane). */
HWPA_HD void hw_tracegen_init_zmw(const hwpa_tracegen_params* p, uint64_t z,
                                  hwpa_tracegen_state* st, int i)
{
    const float us = hw_uniform(hw_hash(p->seed, z, 0, HW_STREAM_SLOPE));
    st->slope[i] = (2.0f * us - 1.0f) * p->drift_per_chunk;
    const float ut = hw_uniform(hw_hash(p->seed, z, 0, HW_STREAM_STEP));
    st->step_frame[i] = hw_floorf(ut * p->step_span);
    const float ue = hw_uniform(hw_hash(p->seed, z, 0, HW_STREAM_EMPTY));
    st->empty[i] = (ue < p->empty_fraction) ? 1 : 0;
    st->in_pulse[i] = 0;
    st->analog[i] = 0;
    st->pad[i] = 0;
}

/* Value of frame t for ZMW z; advances the pulse state. */
HWPA_HD int16_t hw_tracegen_frame(const hwpa_tracegen_params* p, uint64_t z, uint32_t t,
                                  hwpa_tracegen_state* st, int i)
{
    float v = p->baseline_mean + (st->slope[i] * (float)t) / p->drift_chunk;
    if (p->step_drift != 0.0f && (float)t >= st->step_frame[i]) v = v + p->step_drift;
    v = v + p->baseline_sigma * hw_normal(p->seed, z, t, HW_STREAM_NOISE);
    const float ux = hw_uniform(hw_hash(p->seed, z, t, HW_STREAM_EXIT));
    if (st->in_pulse[i]) {
        float amp = p->amplitude[st->analog[i]];
        if (p->snr_hi > 0.0f) amp = amp * hw_tracegen_amp_scale(p, z);
        v = v + amp * (1.0f + p->amplitude_cv * hw_normal(p->seed, z, t, HW_STREAM_AMP));
        if (ux * p->mean_pulse_width < 1.0f) st->in_pulse[i] = 0;
    } else if (!st->empty[i]) {
        if (ux * p->mean_ipd < 1.0f) {
            st->in_pulse[i] = 1;
            const float ua = hw_uniform(hw_hash(p->seed, z, t, HW_STREAM_ANALOG));
            int a = (int)(ua * 4.0f);
            st->analog[i] = (uint8_t)(a > 3 ? 3 : a);
        }
    }
    float r = hw_rintf(v);
    r = hw_minf(hw_maxf(r, -32768.0f), 32767.0f);
    return (int16_t)(int32_t)r;
}

#endif /* HWPA_TRACEGEN_H */
