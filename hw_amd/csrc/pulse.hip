// pulse.hip -- pulse accumulation: segments the frame labels into pulses
// with fixed-point signal statistics and accumulates baseline statistics
// (HostPulseAccumulator, traceAnalysis/HostPulseAccumulator.{h,cpp};
// PulseAccumulator::InitFactory, traceAnalysis/PulseAccumulator.cpp:48-76;
// Pulse fixed point, dataTypes/Pulse.h:212-222).
//
// Mapping: one wave64 per lane, one thread per ZMW; the open segment lives in
// registers for the chunk and in HBM between chunks.  Labels and signal are
// read once (coalesced 128 B rows); pulses go to the ZMW's slot list.

#include "internal.h"

namespace hwpa {

struct SegLane {
    uint32_t start[LS];
    int16_t first[LS];
    int16_t last[LS];
    int16_t smax[LS];
    int16_t label[LS];
    float total[LS];
    float m2[LS];
};

struct PulseConsts {
    float widthThresh, ampThresh;
    uint32_t maxCalls;
    // base code of analog a in byte a.  One word, so a per-ZMW lookup is a shift of a
    // scalar register: a per-lane index into a byte array of the kernel arguments is a
    // memory load, whose wait (loads complete in order) would drain the load ring.
    uint32_t analogMap;
};

struct PulserDev {
    PulseConsts k;
    uint32_t lanes;
    SegLane* seg = nullptr;
};

// std::min / std::max on floats (NaN handling of ToPulse's clamps).
__device__ __forceinline__ float StdMin(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float StdMax(float a, float b) { return (a < b) ? b : a; }
__device__ __forceinline__ uint16_t ToFixed(float v) { return (uint16_t)(v / 0.1f + 0.5f); }

// A finished pulse: the values the record is computed from (HostPulseAccumulator's
// LabelsSegment::ToPulse), kept until the record is written.
struct PulseEnd {
    uint32_t start, stop;
    float meanNum, total, m2;
    int smax, label;
};

// One thread per ZMW.  QUEUE: a finished segment's values go to a per-thread queue in LDS
// (6 words, conflict-free [slot][word][ZMW] layout); when some ZMW's queue could overflow
// during the next batch, the wave writes the queued records with one loop over the longest
// queue.  The record's divisions and fixed-point conversions then run about once per
// pulse of the busiest ZMW instead of at every frame where any ZMW of the wave ends one.
template <int UNROLL, int PD, bool QUEUE>
__global__ __launch_bounds__(64) void PulseKernel(const int16_t* __restrict__ labels, const int16_t* __restrict__ latent,
                                                  const int16_t* __restrict__ traces, int frames, int32_t firstFrame,
                                                  const hwpa_lane_model* __restrict__ models, SegLane* seg,
                                                  hwpa_pulse* __restrict__ pulses, uint32_t* __restrict__ counts,
                                                  uint32_t* __restrict__ overflows, hwpa_stat_acc* __restrict__ blStats,
                                                  PulseConsts k)
{
    // A queue holds QFILL entries plus one batch; it is flushed when any ZMW's holds more
    // than QFILL after a batch, so it never overflows.
    constexpr int QFILL = 4;
    constexpr int QCAP = QUEUE ? UNROLL + QFILL : 1;
    __shared__ uint32_t queue[QCAP * 6 * LS];
    const int lane = blockIdx.x;
    const int z = threadIdx.x;
    SegLane& S = seg[lane];
    uint32_t start = S.start[z];
    int first = S.first[z], last = S.last[z], smax = S.smax[z], label = S.label[z];
    float total = S.total[z], m2 = S.m2[z];
    const float minMean = hw_h2f(models[lane].analogs[NA - 1].means[z]);
    const float maxSignal = 65535 * 0.1f;
    hwpa_pulse* out = pulses + ((size_t)lane * LS + z) * k.maxCalls;
    uint32_t n = 0, ovf = 0, q = 0;
    float bm0 = 0.f, bm1 = 0.f, bm2 = 0.f;
    const int16_t* lb = labels + (size_t)lane * frames * LS + z;
    const int16_t* lt = latent + (size_t)lane * NL * LS + z;
    const int16_t* tr = traces + (size_t)lane * frames * LS + z;

    // LabelsSegment::ToPulse (HostPulseAccumulator.h:64-173) of a finished segment.
    auto emit = [&](const PulseEnd& e) {
        const int width = (int)(e.stop - e.start);
        const float rawMean = e.meanNum / width;
        const float rawMid = e.total / (width - 2);
        const bool keep = (width >= k.widthThresh) || (rawMean * width >= k.ampThresh * minMean);
        int l = e.label;
        if (l > 2 * NA) l -= 2 * NA;
        if (l > NA) l -= NA;
        hwpa_pulse p;
        p.start = e.start;
        p.width = (uint16_t)width;
        p.mean_signal = ToFixed(StdMin(maxSignal, StdMax(0.0f, rawMean)));
        p.mid_signal = ToFixed(width < 3 ? 0.0f : StdMin(maxSignal, StdMax(0.0f, rawMid)));
        p.max_signal = ToFixed(StdMin(maxSignal, StdMax(0.0f, (float)e.smax)));
        p.signal_m2 = e.m2;
        const uint32_t code = (k.analogMap >> (8u * ((uint32_t)(l - 1) & 3u))) & 0xffu;
        p.label = (l >= 1 && l <= NA) ? (uint8_t)code : (uint8_t)4;
        p.is_reject = keep ? 0 : 1;
        p.pad[0] = p.pad[1] = 0;
        if (n < k.maxCalls) out[n++] = p;
        else ++ovf;
    };
    auto qword = [&](uint32_t slot, int w) -> uint32_t& { return queue[(slot * 6 + (uint32_t)w) * LS + z]; };
    auto flush = [&]() {
        for (uint32_t j = 0; __any(j < q); ++j) {
            if (j < q) {
                PulseEnd e;
                e.start = qword(j, 0);
                e.stop = qword(j, 1);
                e.meanNum = __uint_as_float(qword(j, 2));
                e.total = __uint_as_float(qword(j, 3));
                e.m2 = __uint_as_float(qword(j, 4));
                const uint32_t sl = qword(j, 5);
                e.smax = (int)(int16_t)(sl & 0xffffu);
                e.label = (int)(int16_t)(sl >> 16);
                emit(e);
            }
        }
        q = 0;
    };

    // Labels and signal of frames [f0, f0 + UNROLL); frames past the chunk re-read its last
    // frame (never used).  Every load is unconditional, from a selected address: a load under
    // a branch leaves the compiler unable to count the loads issued after an earlier one, and
    // it then waits for all of them (the ring's loads in flight included).
    auto load = [&](int f0, int (&lv)[UNROLL], int (&sv)[UNROLL]) {
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) {
            const int rf = min(f0 + u, frames - 1);
            lv[u] = (int)lb[(size_t)rf * LS];
            const int16_t* sp = (rf < NL) ? lt + (size_t)rf * LS : tr + (size_t)(rf - NL) * LS;
            sv[u] = (int)*sp;
        }
    };
    auto accumulate = [&](int f0, const int (&lv)[UNROLL], const int (&sv)[UNROLL]) {
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) {
            const int rf = f0 + u;
            if (rf >= frames) break;
            const uint32_t absf = (uint32_t)(rf + firstFrame);
            const int lab = lv[u];
            const int sig = sv[u];
            const bool isUp = (lab > NA) && (lab <= 2 * NA);
            const bool boundary = isUp || (lab == 0 && label != 0);
            const bool isPulse = label != 0;
            if (boundary && isPulse) {
                // (float)(total + last + first) of ToPulse, in the same order.
                const float meanNum = (total + last) + first;
                if (QUEUE) {
                    qword(q, 0) = start;
                    qword(q, 1) = absf;
                    qword(q, 2) = __float_as_uint(meanNum);
                    qword(q, 3) = __float_as_uint(total);
                    qword(q, 4) = __float_as_uint(m2);
                    qword(q, 5) = ((uint32_t)smax & 0xffffu) | ((uint32_t)label << 16);
                    ++q;
                } else {
                    emit(PulseEnd{start, absf, meanNum, total, m2, smax, label});
                }
            }
            if (boundary) {
                start = absf;
                first = sig;
                last = 0;
                smax = sig;
                total = 0.0f;
                m2 = 0.0f;
                label = lab;
            } else {
                total = total + last;
                m2 = m2 + (float)last * last;
                last = sig;
                smax = sig > smax ? sig : smax;
            }
            if (!isPulse && !boundary) {
                const float x = (float)sig;
                bm0 += 1.0f;
                bm1 += x;
                bm2 += x * x;
            }
        }
        if (QUEUE && __any(q > (uint32_t)QFILL)) flush();
    };
    // minMean is read before the frame loop (behind the first loads when there is a ring):
    // a value still in flight when the record loop starts would make the compiler wait for
    // every outstanding load before that loop.
    if (PD == 1) {
        asm volatile("" ::"v"(minMean));
        for (int f0 = 0; f0 < frames; f0 += UNROLL) {
            int lv[UNROLL], sv[UNROLL];
            load(f0, lv, sv);
            accumulate(f0, lv, sv);
        }
    } else {
        // A ring of PD batches, each refilled in place with the batch PD batches ahead as
        // soon as it has been accumulated: PD - 1 batches of loads stay in flight, and no
        // register holding an outstanding load is ever copied.
        int lv[PD][UNROLL], sv[PD][UNROLL];
#pragma unroll
        for (int p = 0; p < PD; ++p) load(p * UNROLL, lv[p], sv[p]);
        asm volatile("" ::"v"(minMean));
        for (int f0 = 0; f0 < frames; f0 += PD * UNROLL) {
#pragma unroll
            for (int p = 0; p < PD; ++p) {
                const int fb = f0 + p * UNROLL;
                accumulate(fb, lv[p], sv[p]);
                load(fb + PD * UNROLL, lv[p], sv[p]);
            }
        }
    }
    if (QUEUE) flush();
    S.start[z] = start;
    S.first[z] = (int16_t)first;
    S.last[z] = (int16_t)last;
    S.smax[z] = (int16_t)smax;
    S.label[z] = (int16_t)label;
    S.total[z] = total;
    S.m2[z] = m2;
    counts[(size_t)lane * LS + z] = n;
    overflows[(size_t)lane * LS + z] = ovf;
    hwpa_stat_acc& bs = blStats[lane];
    bs.offset[z] = 0.f;
    bs.m0[z] = bm0;
    bs.m1[z] = bm1;
    bs.m2[z] = bm2;
}

__global__ __launch_bounds__(64) void SegInitKernel(SegLane* seg)
{
    SegLane& S = seg[blockIdx.x];
    const int z = threadIdx.x;
    S.start[z] = 0;
    S.first[z] = 0;
    S.last[z] = 0;
    S.smax[z] = 0;
    S.label[z] = 0;
    S.total[z] = 0.f;
    S.m2[z] = 0.f;
}

int32_t PulserCreate(const hwpa_config& cfg, uint32_t lanes, PulserDev** out)
{
    auto* p = new PulserDev;
    p->lanes = lanes;
    p->k.widthThresh = cfg.xsp_width_thresh;
    const double x = cfg.xsp_amp_thresh;
    p->k.ampThresh = (float)(x / (1.0 - x));
    p->k.maxCalls = cfg.max_calls_per_zmw;
    p->k.analogMap = 0;
    for (int a = 0; a < NA; ++a) {
        const uint8_t b = cfg.movie.analogs[a].base_label;
        const uint32_t code = (b == 'A') ? 0 : (b == 'C') ? 1 : (b == 'G') ? 2 : (b == 'T') ? 3 : 4;
        p->k.analogMap |= code << (8 * a);
    }
    int32_t rc = DeviceAlloc((void**)&p->seg, sizeof(SegLane) * lanes, "pulse segments");
    if (rc == HWPA_OK) {
        hipLaunchKernelGGL(SegInitKernel, dim3(lanes), dim3(LS), 0, nullptr, p->seg);
        rc = CheckLaunch("SegInitKernel");
    }
    if (rc == HWPA_OK && hipDeviceSynchronize() != hipSuccess) rc = Fail(HWPA_ERR_HIP, "pulser init failed");
    if (rc != HWPA_OK) {
        PulserDestroy(p);
        return rc;
    }
    *out = p;
    return HWPA_OK;
}

void PulserDestroy(PulserDev* p)
{
    if (!p) return;
    DeviceFree(p->seg);
    delete p;
}

uint32_t PulserMaxCalls(const PulserDev* p) { return p->k.maxCalls; }

int32_t PulserRun(PulserDev* p, const int16_t* labels, const int16_t* latent, const int16_t* traces, uint32_t frames,
                  int32_t first_frame, const hwpa_lane_model* models, hwpa_pulse* pulses, uint32_t* counts,
                  uint32_t* overflows, hwpa_stat_acc* bl_stats, hipStream_t s)
{
    if (!p || !labels || !latent || !traces || !models || !pulses || !counts || !overflows || !bl_stats)
        return Fail(HWPA_ERR_INVALID_ARG, "pulser_run: null argument");
    if (frames <= (uint32_t)NL) return Fail(HWPA_ERR_INVALID_ARG, "pulser_run: frames must be > 24");
#define HWPA_PULSE_LAUNCH(U, PF, Q)                                                                               \
    hipLaunchKernelGGL((PulseKernel<U, PF, Q>), dim3(p->lanes), dim3(LS), 0, s, labels, latent, traces, (int)frames, \
                       first_frame, models, p->seg, pulses, counts, overflows, bl_stats, p->k)
    // Variants: frames per batch x batches in the load ring x queued record writing.
    switch (TuneGet(kTunePulse, p->lanes, frames)) {
    case 1: HWPA_PULSE_LAUNCH(8, 2, false); break;
    case 2: HWPA_PULSE_LAUNCH(16, 2, false); break;
    case 3: HWPA_PULSE_LAUNCH(8, 1, true); break;
    case 4: HWPA_PULSE_LAUNCH(8, 2, true); break;
    case 5: HWPA_PULSE_LAUNCH(4, 2, true); break;
    case 6: HWPA_PULSE_LAUNCH(4, 1, true); break;
    case 7: HWPA_PULSE_LAUNCH(4, 4, true); break;
    case 8: HWPA_PULSE_LAUNCH(8, 3, true); break;
    case 9: HWPA_PULSE_LAUNCH(4, 8, true); break;
    case 10: HWPA_PULSE_LAUNCH(4, 12, true); break;
    case 11: HWPA_PULSE_LAUNCH(8, 4, true); break;
    case 12: HWPA_PULSE_LAUNCH(2, 16, true); break;
    default: HWPA_PULSE_LAUNCH(8, 1, false); break;
    }
#undef HWPA_PULSE_LAUNCH
    return CheckLaunch("PulseKernel");
}

}  // namespace hwpa
