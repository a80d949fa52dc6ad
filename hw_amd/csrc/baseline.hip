// baseline.hip -- multi-scale erode/dilate baseline filter, baseline
// subtraction and per-chunk baseline statistics on gfx950.
//
// Semantics follow HostMultiScaleBaseliner (traceAnalysis/HostMultiScaleBaseliner.cpp:57-212):
// lower envelope = opening cascade, upper = closing then openings; every
// extremum filter returns op(previous W inputs) of its (decimated) stream
// (traceAnalysis/TraceFilters.cpp:10-57), state carried across chunks.  Any
// algorithm that returns that window extremum gives the reference's bits.
//
// Mapping: one wave64 per lane, one thread per ZMW, frames streamed in
// groups of the aggregate stride (16).  The group's raw frames are read once
// into registers, pushed through both cascades, and the same registers feed
// the subtraction and the statistics, so HBM sees the trace once in and once
// out.
//
// Filters run on packed 16-bit pairs so one v_pk_min_i16 / v_pk_max_i16
// serves two filters: stage 0 pairs (x, ~x) -- bitwise NOT reverses the
// order, so the upper cascade's dilate-then-erode becomes erode-then-dilate
// of ~x -- and later stages pair (lower, upper), which both run
// erode-then-dilate.  The non-final stages keep the last W inputs of each
// filter pair in registers and compute a group's window extrema with an
// unrolled block prefix/suffix pass (van Herk / Gil-Werman with blocks of W
// over history ++ group), so they need no ring and no LDS.  The final stage,
// one sample per group, keeps a van Herk / Gil-Werman ring per filter pair in
// HBM whose lane-uniform head is shared by both pairs.

#include "internal.h"

#include <algorithm>
#include <cmath>

namespace hwpa {

typedef short s16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint32_t PMin(uint32_t a, uint32_t b)
{
    return __builtin_bit_cast(uint32_t,
                              __builtin_elementwise_min(__builtin_bit_cast(s16x2, a), __builtin_bit_cast(s16x2, b)));
}
__device__ __forceinline__ uint32_t PMax(uint32_t a, uint32_t b)
{
    return __builtin_bit_cast(uint32_t,
                              __builtin_elementwise_max(__builtin_bit_cast(s16x2, a), __builtin_bit_cast(s16x2, b)));
}
template <bool MX>
__device__ __forceinline__ uint32_t POp(uint32_t a, uint32_t b) { return MX ? PMax(a, b) : PMin(a, b); }
__device__ __forceinline__ uint32_t Pack(int lo, int hi) { return (uint32_t)(uint16_t)lo | ((uint32_t)(uint16_t)hi << 16); }
__device__ __forceinline__ int LoHalf(uint32_t p) { return (int)(int16_t)(uint16_t)(p & 0xffffu); }
__device__ __forceinline__ int HiHalf(uint32_t p) { return (int)(int16_t)(uint16_t)(p >> 16); }
constexpr uint32_t HI_NOT = 0xffff0000u;

struct BlLaneState {
    float ema_sum[LS];
    float ema_weight[LS];
    float sigma_ema[LS];
    int16_t lat_data[LS];
    int16_t lat_raw[LS];
    uint8_t lat_l[LS];
    uint8_t lat_h1[LS];
    uint32_t hold[2][LS];  // final-stage van Herk / Gil-Werman hold, packed (erode pair, dilate pair)
    uint32_t last[2][LS];
    int32_t counter[LS];   // [0]: final-stage ring head (lane-uniform)
};

struct BlConsts {
    float cMeanBias, cSigmaBias, meanEmaAlpha, sigmaEmaAlpha, jumpTol, scaler;
    int32_t pedestal;
};

struct BaselinerDev {
    hwpa_filter_params fp;
    BlConsts k;
    uint32_t lanes;
    size_t winPerLane;  // packed u32 rows per lane
    BlLaneState* state = nullptr;
    uint32_t* win = nullptr;
};

// out[k] = op(c[k .. k+W-1]) for c = hist ++ in, k < N; then hist := c[N .. N+W-1].
// Blocks of W start at index 0 of c: a window either is one block or spans
// the suffix of one block and the prefix of the next.  Prefixes are formed
// forward for c[W ..], suffixes backward and combined as they are formed.
template <int W, int N, bool MX>
__device__ __forceinline__ void Slide(uint32_t (&hist)[W], const uint32_t (&in)[N], uint32_t (&out)[N])
{
    constexpr int L = W + N;
    constexpr int E = ((N - 1) / W) * W + W - 1;  // end of the block holding index N-1 (< L - 1)
    uint32_t c[L], P[L];
#pragma unroll
    for (int i = 0; i < W; ++i) c[i] = hist[i];
#pragma unroll
    for (int i = 0; i < N; ++i) c[W + i] = in[i];
#pragma unroll
    for (int i = W; i < L; ++i) {
        if (i % W == 0)
            P[i] = c[i];
        else
            P[i] = POp<MX>(P[i - 1], c[i]);
    }
    uint32_t run = 0;
#pragma unroll
    for (int i = E; i >= 0; --i) {
        if (i % W == W - 1)
            run = c[i];
        else
            run = POp<MX>(c[i], run);
        if (i < N) out[i] = (i % W == 0) ? run : POp<MX>(run, P[i + W - 1]);
    }
#pragma unroll
    for (int i = 0; i < W; ++i) hist[i] = c[N + i];
}

template <int NST, int S0, int S1, int S2, int W0, int W1, int W2>
struct Table {
    static constexpr int nst = NST;
    static constexpr int s[3] = {S0, S1, S2};
    static constexpr int w[3] = {W0, W1, W2};
    static constexpr int S = S0 * S1 * (NST == 3 ? S2 : 1);
    static constexpr int N0 = S / S0;                       // stage-0 samples per group
    static constexpr int N1 = (NST == 3) ? N0 / S1 : 1;     // stage-1 samples per group
    static constexpr int WF = (NST == 3) ? W2 : W1;         // final-stage width
    static constexpr int W1R = (NST == 3) ? W1 : 1;         // register stage-1 width (unused for 2 stages)
    // packed rows per lane: stage-0 histories, stage-1 histories, final rings
    static constexpr int rowsHist = 2 * W0 + (NST == 3 ? 2 * W1 : 0);
    static constexpr int winPerLane = rowsHist + 2 * WF;
};

// One van Herk / Gil-Werman step of the final stage on a packed ring (HBM).
// `front` is ring[ctr] read at the start of the group (after any suffix refill).
template <bool MX>
__device__ __forceinline__ uint32_t RingResult(int ctr, uint32_t front, uint32_t& s, uint32_t last)
{
    s = (ctr == 0) ? last : POp<MX>(s, last);
    return POp<MX>(s, front);
}

// The rings now hold one complete block each: replace each by its suffix extrema (erode
// pair ringA, dilate pair ringB), working down from the top in batches of 12 loads per ring;
// both rings' batches are in flight together.
template <int W>
__device__ __forceinline__ void RingRefill2(uint32_t* ringA, uint32_t lastA, uint32_t* ringB, uint32_t lastB)
{
    constexpr int B = 12;
    uint32_t accA = lastA, accB = lastB;
#pragma unroll 1
    for (int top = W - 2; top >= 0; top -= B) {
        uint32_t va[B], vb[B];
#pragma unroll
        for (int q = 0; q < B; ++q) {
            va[q] = (top - q >= 0) ? ringA[(top - q) * LS] : 0u;
            vb[q] = (top - q >= 0) ? ringB[(top - q) * LS] : 0u;
        }
#pragma unroll
        for (int q = 0; q < B; ++q) {
            if (top - q >= 0) {
                accA = PMin(accA, va[q]);
                ringA[(top - q) * LS] = accA;
                accB = PMax(accB, vb[q]);
                ringB[(top - q) * LS] = accB;
            }
        }
    }
}

// The raw samples of a group of S frames, two 16-bit values per register (int16 samples,
// or uint8 zero-extended: both fit a short), loaded straight into the halves.
typedef short Pair16 __attribute__((ext_vector_type(2)));
template <int S>
struct Group16 {
    static_assert(S % 2 == 0, "whole pairs");
    Pair16 p[S / 2];
    __device__ __forceinline__ int At(int j) const { return (j & 1) ? (int)p[j >> 1].y : (int)p[j >> 1].x; }
};

// Loads a group's samples of one ZMW column (src: its first frame); the raw trace is read once.
template <int S, class E>
__device__ __forceinline__ void LoadGroup(Group16<S>& g, const E* src)
{
#pragma unroll
    for (int k = 0; k < S / 2; ++k) {
        g.p[k].x = (short)__builtin_nontemporal_load(&src[(size_t)(2 * k) * LS]);
        g.p[k].y = (short)__builtin_nontemporal_load(&src[(size_t)(2 * k + 1) * LS]);
    }
}

// E: raw sample type (int16_t, or uint8_t zero-extended as BatchData::Extract does,
// dataTypes/BatchData.h:160-182); LPB lanes (waves) per workgroup -- the autotuned lane
// tile; PD groups of raw frames in flight (1 or 2).
template <class T, class E, int LPB, int PD>
__global__ __launch_bounds__(64 * LPB) void BaselineKernel(const E* __restrict__ raw, int frames, int lanes,
                                                           int16_t* __restrict__ out,
                                                           hwpa_baseliner_stats* __restrict__ stats,
                                                           BlLaneState* __restrict__ state,
                                                           uint32_t* __restrict__ win, BlConsts k)
{
    constexpr int S = T::S;
    constexpr int NST = T::nst;
    constexpr int W0 = T::w[0], N0 = T::N0, N1 = T::N1, W1 = T::W1R, WF = T::WF;

    const int lane = blockIdx.x * LPB + (int)(threadIdx.x >> 6);
    const int z = threadIdx.x & 63;
    if (lane >= lanes) return;
    BlLaneState& st = state[lane];
    uint32_t* wl = win + (size_t)lane * T::winPerLane * LS + z;

    uint32_t h0a[W0], h0b[W0];
#pragma unroll
    for (int i = 0; i < W0; ++i) {
        h0a[i] = wl[i * LS];
        h0b[i] = wl[(W0 + i) * LS];
    }
    // Stage-1 windows wait in LDS between groups (each thread its own column): they are used
    // once per group, and their 2 W1 registers are free for the rest of it.
    __shared__ uint32_t sH1[LPB][2 * W1][LS];
    uint32_t(*h1)[LS] = sH1[threadIdx.x >> 6];
    if constexpr (NST == 3) {
#pragma unroll
        for (int i = 0; i < 2 * W1; ++i) h1[i][z] = wl[(2 * W0 + i) * LS];
    }
    uint32_t* ringA = wl + T::rowsHist * LS;
    uint32_t* ringB = ringA + WF * LS;
    uint32_t sA = st.hold[0][z], sB = st.hold[1][z];
    uint32_t lastA = st.last[0][z], lastB = st.last[1][z];
    int ctr = st.counter[0];

    float emaSum = st.ema_sum[z], emaWeight = st.ema_weight[z], sigmaEma = st.sigma_ema[z];
    int latData = st.lat_data[z], latRaw = st.lat_raw[z];
    bool latL = st.lat_l[z] != 0, latH1 = st.lat_h1[z] != 0;

    // per-chunk statistics (offsets are 0)
    // (the sample and group counts are the chunk's frame and group counts; the first four
    // samples go straight to the output record)
    hwpa_baseliner_stats& o = stats[lane];
    float cm1 = 0.f, cm2 = 0.f, lagM2 = 0.f;
    float bbuf[4] = {0.f, 0.f, 0.f, 0.f};
    int tmin = 32767, tmax = -32768;
    float bm0 = 0.f, bm1 = 0.f, bm2 = 0.f, rawBl = 0.f;
    float gm1 = 0.f, gm2 = 0.f;

    const E* rl = raw + (size_t)lane * frames * LS + z;
    int16_t* ol = out + (size_t)lane * frames * LS + z;
    const int groups = frames / S;
    const float minSigma = .288675135f;

    static_assert(PD == 1 || PD == 2, "prefetch depth");
    Group16<S> rv, pv;
    LoadGroup(rv, rl);
    if (PD == 2) LoadGroup(pv, rl + (size_t)((groups > 1) ? 1 : 0) * S * LS);

    for (int g = 0; g < groups; ++g) {
        // Final-stage ring: refill on wrap, then read the fronts early.
        if (ctr == 0) {
            RingRefill2<WF>(ringA, lastA, ringB, lastB);
        }
        const uint32_t frontA = ringA[ctr * LS];
        const uint32_t frontB = ringB[ctr * LS];
        // Prefetch the raw frames of group g + PD (the last group re-read past the end).
        Group16<S> nv;
        const int gn = (g + PD < groups) ? g + PD : groups - 1;
        LoadGroup(nv, rl + (size_t)gn * S * LS);

        // Stage 0 on (x, ~x): erode pair, then dilate pair -> (lower, ~upper), in two halves of
        // the group (the same window extrema, with about half the live registers).
        constexpr int H0 = N0 / 2;
        static_assert(N0 % 2 == 0 && (NST == 2 || H0 % T::s[1] == 0), "stage-1 inputs fall in whole halves");
        uint32_t a1[N1];
        uint32_t fin = 0;
#pragma unroll
        for (int hf = 0; hf < 2; ++hf) {
            uint32_t a0[H0], e0[H0], d0[H0];
#pragma unroll
            for (int i = 0; i < H0; ++i) {
                const int x = rv.At((hf * H0 + i) * T::s[0]);
                a0[i] = Pack(x, ~x);
            }
            Slide<W0, H0, false>(h0a, a0, e0);
            Slide<W0, H0, true>(h0b, e0, d0);
            if constexpr (NST == 3) {
#pragma unroll
                for (int i = 0; i < H0 / T::s[1]; ++i) a1[hf * (H0 / T::s[1]) + i] = d0[i * T::s[1]] ^ HI_NOT;
            } else {
                if (hf == 0) fin = d0[0] ^ HI_NOT;
            }
        }
        if constexpr (NST == 3) {
            uint32_t h1a[W1], h1b[W1], e1[N1], d1[N1];
#pragma unroll
            for (int i = 0; i < W1; ++i) {
                h1a[i] = h1[i][z];
                h1b[i] = h1[W1 + i][z];
            }
            Slide<W1, N1, false>(h1a, a1, e1);
            Slide<W1, N1, true>(h1b, e1, d1);
#pragma unroll
            for (int i = 0; i < W1; ++i) {
                h1[i][z] = h1a[i];
                h1[W1 + i][z] = h1b[i];
            }
            // Keeps the windows in LDS across groups (no promotion back to registers).
            asm volatile("" ::: "memory");
            fin = d1[0];
        }
        // Final stage: erode pair then dilate pair.
        const uint32_t fe = RingResult<false>(ctr, frontA, sA, lastA);
        ringA[ctr * LS] = fin;
        lastA = fin;
        const uint32_t fd = RingResult<true>(ctr, frontB, sB, lastB);
        ringB[ctr * LS] = fe;
        lastB = fe;
        ctr = (ctr + 1 == WF) ? 0 : ctr + 1;
        const int lo = LoHalf(fd), up = HiHalf(fd);

        // GetSmoothedBlEstimate (HostMultiScaleBaseliner.cpp:178-212)
        const int16_t lv = hw_wrap16(lo - k.pedestal);
        const int16_t uv = hw_wrap16(up - k.pedestal);
        const float sigma = hw_maxf((float)hw_wrap16((int32_t)uv - (int32_t)lv) / k.cSigmaBias, minSigma);
        const float newSigmaEma = k.sigmaEmaAlpha * sigmaEma + (1.0f - k.sigmaEmaAlpha) * sigma;
        const float blEst = 0.5f * (float)hw_wrap16((int32_t)uv + (int32_t)lv) + k.cMeanBias * newSigmaEma;
        const bool mask = (emaWeight == 0.0f) || (blEst - emaSum / emaWeight < k.jumpTol * sigmaEma);
        const float newWeight = k.meanEmaAlpha * emaWeight + (1.0f - k.meanEmaAlpha);
        const float newSum = k.meanEmaAlpha * emaSum + (1.0f - k.meanEmaAlpha) * blEst;
        if (mask) {
            emaWeight = newWeight;
            emaSum = newSum;
            sigmaEma = newSigmaEma;
        }
        const float subBl = (emaSum / emaWeight) * k.scaler;

        // Subtraction and statistics (HostMultiScaleBaseliner.cpp:124-176)
        const float thr = (4.5f * sigmaEma) * k.scaler;
#pragma unroll
        for (int j = 0; j < S; ++j) {
            const float rawScaled = (float)hw_wrap16(rv.At(j) - k.pedestal) * k.scaler;
            const int bls = hw_round_i16(rawScaled - subBl);
            (ol + (size_t)g * S * LS)[j * LS] = (int16_t)bls;   // cached: the histogram reads it next
            const int rr = hw_round_i16(rawScaled);
            const bool maskHp1 = (float)bls < thr;
            const bool m = latH1 && latL && maskHp1;
            latL = (float)bls < thr;
            latH1 = maskHp1;
            // BaselinerStatAccumulator::AddSample on the latent frame
            const float bs = (float)latData;
            if (g == 0 && j < 4) o.full.fbuf[j][z] = bs;
            lagM2 += bbuf[j & 3] * bs;
            bbuf[j & 3] = bs;
            cm1 += bs;
            cm2 += bs * bs;
            tmin = latData < tmin ? latData : tmin;
            tmax = latData > tmax ? latData : tmax;
            // Branch-free masked sums: adding +0 leaves a sum of integers unchanged.
            bm0 = bm0 + (m ? 1.0f : 0.0f);
            bm1 = bm1 + (m ? bs : 0.0f);
            bm2 = bm2 + (m ? bs * bs : 0.0f);
            rawBl += (float)(m ? latRaw : 0);
            latRaw = rr;
            latData = bls;
        }
        const float bgv = (float)hw_round_i16(subBl);
        gm1 += bgv;
        gm2 += bgv * bgv;
        rv = (PD == 2) ? pv : nv;
        pv = nv;
    }

    // Persist carried state.
#pragma unroll
    for (int i = 0; i < W0; ++i) {
        wl[i * LS] = h0a[i];
        wl[(W0 + i) * LS] = h0b[i];
    }
    if constexpr (NST == 3) {
#pragma unroll
        for (int i = 0; i < 2 * W1; ++i) wl[(2 * W0 + i) * LS] = h1[i][z];
    }
    st.hold[0][z] = sA;
    st.hold[1][z] = sB;
    st.last[0][z] = lastA;
    st.last[1][z] = lastB;
    if (z == 0) st.counter[0] = ctr;
    st.ema_sum[z] = emaSum;
    st.ema_weight[z] = emaWeight;
    st.sigma_ema[z] = sigmaEma;
    st.lat_data[z] = (int16_t)latData;
    st.lat_raw[z] = (int16_t)latRaw;
    st.lat_l[z] = latL ? 1 : 0;
    st.lat_h1[z] = latH1 ? 1 : 0;

    o.full.basic.offset[z] = 0.f;
    o.full.basic.m0[z] = (float)frames;   // exact: frames < 2^24
    o.full.basic.m1[z] = cm1;
    o.full.basic.m2[z] = cm2;
    o.full.lag_m2[z] = lagM2;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        o.full.bbuf[q][z] = bbuf[q];
    }
    o.full.bidx[0][z] = (uint8_t)(frames < 4 ? frames : 4);
    o.full.bidx[1][z] = (uint8_t)(frames & 3);
    o.trace_min[z] = (int16_t)tmin;
    o.trace_max[z] = (int16_t)tmax;
    o.baseline.offset[z] = 0.f;
    o.baseline.m0[z] = bm0;
    o.baseline.m1[z] = bm1;
    o.baseline.m2[z] = bm2;
    o.raw_baseline_sum[z] = rawBl;
    o.background.offset[z] = 0.f;
    o.background.m0[z] = (float)groups;
    o.background.m1[z] = gm1;
    o.background.m2[z] = gm2;
}

// Initial state: every filter window holds W zeros (traceAnalysis/BlockFilterStage.h:28-31);
// in the stage-0 (x, ~x) domain the upper half of a zero window is ~0.
__global__ void BaselineInitKernel(BlLaneState* state, uint32_t* win, int winPerLane, int stage0Rows)
{
    const int lane = blockIdx.x;
    const int z = threadIdx.x;
    BlLaneState& st = state[lane];
    st.ema_sum[z] = 0.f;
    st.ema_weight[z] = 0.f;
    st.sigma_ema[z] = 2.0f;
    st.lat_data[z] = 0;
    st.lat_raw[z] = 0;
    st.lat_l[z] = 0;
    st.lat_h1[z] = 0;
    for (int f = 0; f < 2; ++f) {
        st.hold[f][z] = 0;
        st.last[f][z] = 0;
    }
    st.counter[z] = 0;
    uint32_t* wl = win + (size_t)lane * winPerLane * LS;
    for (int i = 0; i < winPerLane; ++i) wl[i * LS + z] = (i < stage0Rows) ? HI_NOT : 0u;
}

using TMSL = Table<3, 1, 2, 8, 11, 17, 61>;
using TMSM = Table<3, 1, 2, 8, 9, 17, 31>;
using TMSS = Table<3, 1, 2, 8, 7, 17, 17>;
using TTSL = Table<2, 2, 8, 1, 11, 61, 1>;
using TTSM = Table<2, 2, 8, 1, 9, 31, 1>;
using TTSS = Table<2, 2, 8, 1, 7, 17, 1>;

template <class T>
static int32_t Launch(BaselinerDev* b, const void* raw, int32_t encoding, uint32_t frames, int16_t* out,
                      hwpa_baseliner_stats* stats, hipStream_t s)
{
    if (frames % T::S != 0) return Fail(HWPA_ERR_INVALID_ARG, "frames must be a multiple of the aggregate stride");
    if ((size_t)T::winPerLane != b->winPerLane) return Fail(HWPA_ERR_UNSUPPORTED, "baseliner: window layout mismatch");
    // Variants: lanes per workgroup (1, 2, 4) x raw-frame groups in flight (1, 2).
    const int v = TuneGet(kTuneBaseline, b->lanes, frames);
    const int lpb = (v % 3 == 1) ? 2 : (v % 3 == 2 ? 4 : 1);
    const bool deep = v >= 3;
    const unsigned grid = (b->lanes + (unsigned)lpb - 1) / (unsigned)lpb;
#define HWPA_BL_LAUNCH(L, P)                                                                                       \
    do {                                                                                                           \
        if (encoding == HWPA_ENCODING_UINT8)                                                                       \
            hipLaunchKernelGGL((BaselineKernel<T, uint8_t, L, P>), dim3(grid), dim3(64 * L), 0, s,                 \
                               (const uint8_t*)raw, (int)frames, (int)b->lanes, out, stats, b->state, b->win, b->k); \
        else                                                                                                       \
            hipLaunchKernelGGL((BaselineKernel<T, int16_t, L, P>), dim3(grid), dim3(64 * L), 0, s,                 \
                               (const int16_t*)raw, (int)frames, (int)b->lanes, out, stats, b->state, b->win, b->k); \
    } while (0)
    if (lpb == 1) {
        if (deep) HWPA_BL_LAUNCH(1, 2); else HWPA_BL_LAUNCH(1, 1);
    } else if (lpb == 2) {
        if (deep) HWPA_BL_LAUNCH(2, 2); else HWPA_BL_LAUNCH(2, 1);
    } else {
        if (deep) HWPA_BL_LAUNCH(4, 2); else HWPA_BL_LAUNCH(4, 1);
    }
#undef HWPA_BL_LAUNCH
    return CheckLaunch("BaselineKernel");
}

uint32_t BaselinerStartupLatency(const hwpa_config& cfg)
{
    // MultiScaleBaseliner::StartupLatency (MultiScaleBaseliner.h:56-99)
    const hwpa_filter_params fp = FilterParams(cfg.filter_type);
    const float meanAlpha = std::pow(0.5f, 1.0f / cfg.mean_ema_scale_strides);
    const float sigmaAlpha = std::exp2(-1.0f / cfg.sigma_ema_scale_strides);
    const float mess = -1.0f / std::log2(meanAlpha);
    const float sess = -1.0f / std::log2(sigmaAlpha);
    return (uint32_t)((float)fp.latent_size + std::max(mess, sess) * (float)fp.aggregate_stride);
}

int32_t BaselinerCreate(const hwpa_config& cfg, uint32_t lanes, BaselinerDev** out)
{
    auto* b = new BaselinerDev;
    b->fp = FilterParams(cfg.filter_type);
    b->lanes = lanes;
    // LaneBaseliner ctor (HostMultiScaleBaseliner.h:58-66), MultiScaleBaseliner::Configure
    // (MultiScaleBaseliner.cpp:53-75)
    // Host constants from libm's float functions, as the reference computes them.
    b->k.cMeanBias = b->fp.mean_bias * std::exp2(cfg.mean_bias_adjust);
    b->k.cSigmaBias = b->fp.sigma_bias * std::exp2(cfg.sigma_bias_adjust);
    b->k.meanEmaAlpha = std::pow(0.5f, 1.0f / cfg.mean_ema_scale_strides);
    b->k.sigmaEmaAlpha = std::exp2(-1.0f / cfg.sigma_ema_scale_strides);
    b->k.jumpTol = (cfg.jump_suppression > 0.0f) ? 1.0f / cfg.jump_suppression : INFINITY;
    b->k.scaler = cfg.movie.photoelectron_sensitivity;
    b->k.pedestal = cfg.movie.pedestal;
    // Packed (two filters per u32) windows: 2 rows per width of every stage.
    size_t w = 0;
    for (int i = 0; i < b->fp.num_stages; ++i) w += 2 * (size_t)b->fp.widths[i];
    b->winPerLane = w;
    int32_t rc = DeviceAlloc((void**)&b->state, sizeof(BlLaneState) * lanes, "baseliner state");
    if (rc == HWPA_OK) rc = DeviceAlloc((void**)&b->win, w * LS * sizeof(uint32_t) * lanes, "baseliner windows");
    if (rc != HWPA_OK) {
        BaselinerDestroy(b);
        return rc;
    }
    hipLaunchKernelGGL(BaselineInitKernel, dim3(lanes), dim3(LS), 0, nullptr, b->state, b->win, (int)w,
                       2 * b->fp.widths[0]);
    rc = CheckLaunch("BaselineInitKernel");
    if (rc == HWPA_OK && hipDeviceSynchronize() != hipSuccess) rc = Fail(HWPA_ERR_HIP, "baseliner init failed");
    if (rc != HWPA_OK) {
        BaselinerDestroy(b);
        return rc;
    }
    *out = b;
    return HWPA_OK;
}

void BaselinerDestroy(BaselinerDev* b)
{
    if (!b) return;
    DeviceFree(b->state);
    DeviceFree(b->win);
    delete b;
}

int32_t BaselinerRun(BaselinerDev* b, const void* raw, uint32_t frames, int16_t* out,
                     hwpa_baseliner_stats* stats, hipStream_t s, int32_t encoding)
{
    if (!b || !raw || !out || !stats || frames == 0) return Fail(HWPA_ERR_INVALID_ARG, "baseliner: bad argument");
    if (encoding != HWPA_ENCODING_INT16 && encoding != HWPA_ENCODING_UINT8)
        return Fail(HWPA_ERR_INVALID_ARG, "Unexpected data encoding in baseline filter");
    const int32_t e = encoding;
    switch (b->fp.num_stages == 3 ? b->fp.widths[0] * 100 + b->fp.widths[2] : 1000 + b->fp.widths[0] * 100 + b->fp.widths[1]) {
    case 1161: return Launch<TMSL>(b, raw, e, frames, out, stats, s);
    case 931: return Launch<TMSM>(b, raw, e, frames, out, stats, s);
    case 717: return Launch<TMSS>(b, raw, e, frames, out, stats, s);
    case 2161: return Launch<TTSL>(b, raw, e, frames, out, stats, s);
    case 1931: return Launch<TTSM>(b, raw, e, frames, out, stats, s);
    case 1717: return Launch<TTSS>(b, raw, e, frames, out, stats, s);
    default: return Fail(HWPA_ERR_UNSUPPORTED, "baseliner: unsupported filter table");
    }
}

}  // namespace hwpa
