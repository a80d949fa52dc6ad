// baseline.hip -- multi-scale erode/dilate baseline filter, baseline
// subtraction and per-chunk baseline statistics on gfx950.
//
// Semantics follow HostMultiScaleBaseliner (traceAnalysis/HostMultiScaleBaseliner.cpp:57-212):
// lower envelope = opening cascade, upper = closing then openings; every
// extremum filter returns op(previous W inputs) of its (decimated) stream
// (traceAnalysis/TraceFilters.cpp:10-57), state carried across chunks.
//
// Mapping: one wave64 per lane, one lane (thread) per ZMW, frames streamed
// in groups of the aggregate stride (16): the group's raw frames are read
// once into registers, pushed through both cascades, and the same registers
// feed the subtraction and the statistics, so HBM sees the trace once in and
// once out.  The van Herk/Gil-Werman ring of every non-final stage lives in
// LDS for the duration of the chunk (one shared, lane-uniform counter per
// stage, since all four filters of a stage advance together); the final
// stage runs once per group and keeps its ring directly in HBM.

#include "internal.h"

#include <algorithm>
#include <cmath>

namespace hwpa {

struct BlLaneState {
    float ema_sum[LS];
    float ema_weight[LS];
    float sigma_ema[LS];
    int16_t lat_data[LS];
    int16_t lat_raw[LS];
    uint8_t lat_l[LS];
    uint8_t lat_h1[LS];
    int16_t hold[4 * HWPA_MAX_FILTER_STAGES][LS];
    int16_t last[4 * HWPA_MAX_FILTER_STAGES][LS];
    int32_t counter[LS];  // [0..stages) used; lane-uniform
};

struct BlConsts {
    float cMeanBias, cSigmaBias, meanEmaAlpha, sigmaEmaAlpha, jumpTol, scaler;
    int32_t pedestal;
};

struct BaselinerDev {
    hwpa_filter_params fp;
    BlConsts k;
    uint32_t lanes;
    size_t winPerLane;  // int16 ring entries per ZMW column
    BlLaneState* state = nullptr;
    int16_t* win = nullptr;
};

template <bool MX>
__device__ __forceinline__ int Ext(int a, int b) { return MX ? (a > b ? a : b) : (a < b ? a : b); }

// One van Herk / Gil-Werman step on a ring of W int16 (ZMW-minor, stride LS).
// ctr is the lane-uniform ring head; when it wraps to 0 the ring (now one
// complete block) is replaced in place by its suffix extrema.
template <int W, bool MX, typename P>
__device__ __forceinline__ int HgwStep(P ring, int ctr, int z, int x, int& s, int& last)
{
    if (ctr == 0) {
        int v = last;
#pragma unroll
        for (int j = W - 2; j >= 0; --j) {
            v = Ext<MX>(v, (int)ring[j * LS + z]);
            ring[j * LS + z] = (int16_t)v;
        }
        s = last;
    } else {
        s = Ext<MX>(s, last);
    }
    const int front = ring[ctr * LS + z];
    ring[ctr * LS + z] = (int16_t)x;
    last = x;
    return Ext<MX>(s, front);
}

template <int NST, int S0, int S1, int S2, int W0, int W1, int W2>
struct Table {
    static constexpr int nst = NST;
    static constexpr int s[3] = {S0, S1, S2};
    static constexpr int w[3] = {W0, W1, W2};
    static constexpr int S = S0 * S1 * (NST == 3 ? S2 : 1);
    static constexpr int ldsW = (NST == 3) ? (W0 + W1) : W0;  // widths of the LDS stages
    static constexpr int lastW = (NST == 3) ? W2 : W1;
    static constexpr int off(int i) { return i == 0 ? 0 : (i == 1 ? 4 * W0 : 4 * (W0 + W1)); }
    static constexpr int winPerLane = 4 * (W0 + W1 + (NST == 3 ? W2 : 0));
};

// Runs stage I of the cascade on n consumed samples in[] -> lo[], up[].
template <class T, int I, int N, typename P>
__device__ __forceinline__ void RunStage(P base, int z, int& ctr, const int* inLo, const int* inUp, int stepIn,
                                         int* outLo, int* outUp, int (&s)[12], int (&last)[12])
{
    constexpr int W = T::w[I];
    constexpr bool upperFirstMax = (I == 0);
    P rLE = base + 0 * W * LS;
    P rLD = base + 1 * W * LS;
    P rU1 = base + 2 * W * LS;
    P rU2 = base + 3 * W * LS;
#pragma unroll
    for (int k = 0; k < N; ++k) {
        const int xl = inLo[k * stepIn];
        const int xu = inUp[k * stepIn];
        const int a = HgwStep<W, false>(rLE, ctr, z, xl, s[4 * I + 0], last[4 * I + 0]);
        outLo[k] = HgwStep<W, true>(rLD, ctr, z, a, s[4 * I + 1], last[4 * I + 1]);
        const int b = HgwStep<W, upperFirstMax>(rU1, ctr, z, xu, s[4 * I + 2], last[4 * I + 2]);
        outUp[k] = HgwStep<W, !upperFirstMax>(rU2, ctr, z, b, s[4 * I + 3], last[4 * I + 3]);
        ctr = (ctr + 1 == W) ? 0 : ctr + 1;
    }
}

template <class T>
__global__ __launch_bounds__(64) void BaselineKernel(const int16_t* __restrict__ raw, int frames,
                                                     int16_t* __restrict__ out,
                                                     hwpa_baseliner_stats* __restrict__ stats,
                                                     BlLaneState* __restrict__ state,
                                                     int16_t* __restrict__ win, BlConsts k)
{
    constexpr int S = T::S;
    constexpr int NST = T::nst;
    __shared__ int16_t lds[4 * T::ldsW * LS];

    const int lane = blockIdx.x;
    const int z = threadIdx.x;
    BlLaneState& st = state[lane];
    int16_t* wl = win + (size_t)lane * T::winPerLane * LS;

    // Load the LDS rings of the non-final stages.
#pragma unroll 4
    for (int i = 0; i < 4 * T::ldsW; ++i) lds[i * LS + z] = wl[i * LS + z];
    int16_t* lastRing = wl + T::off(NST - 1) * LS;

    int s[12], last[12];
#pragma unroll
    for (int f = 0; f < 12; ++f) {
        s[f] = (f < 4 * NST) ? (int)st.hold[f][z] : 0;
        last[f] = (f < 4 * NST) ? (int)st.last[f][z] : 0;
    }
    int ctr0 = st.counter[0];
    int ctr1 = st.counter[1];
    int ctr2 = (NST == 3) ? st.counter[2] : 0;

    float emaSum = st.ema_sum[z], emaWeight = st.ema_weight[z], sigmaEma = st.sigma_ema[z];
    int latData = st.lat_data[z], latRaw = st.lat_raw[z];
    bool latL = st.lat_l[z] != 0, latH1 = st.lat_h1[z] != 0;
    __syncthreads();

    // per-chunk statistics (offsets are 0)
    float cm0 = 0.f, cm1 = 0.f, cm2 = 0.f, lagM2 = 0.f;
    float fbuf[4] = {0.f, 0.f, 0.f, 0.f}, bbuf[4] = {0.f, 0.f, 0.f, 0.f};
    int tmin = 32767, tmax = -32768;
    float bm0 = 0.f, bm1 = 0.f, bm2 = 0.f, rawBl = 0.f;
    float gm0 = 0.f, gm1 = 0.f, gm2 = 0.f;

    const int16_t* rl = raw + (size_t)lane * frames * LS + z;
    int16_t* ol = out + (size_t)lane * frames * LS + z;
    const int groups = frames / S;
    const float minSigma = .288675135f;

    for (int g = 0; g < groups; ++g) {
        int rv[S];
#pragma unroll
        for (int j = 0; j < S; ++j) rv[j] = rl[(size_t)(g * S + j) * LS];

        int lo, up;
        if constexpr (NST == 3) {
            constexpr int n0 = S / T::s[0];
            constexpr int n1 = n0 / T::s[1];
            int lo0[n0], up0[n0], lo1[n1], up1[n1];
            RunStage<T, 0, n0>(lds, z, ctr0, rv, rv, T::s[0], lo0, up0, s, last);
            RunStage<T, 1, n1>(lds + 4 * T::w[0] * LS, z, ctr1, lo0, up0, T::s[1], lo1, up1, s, last);
            int lo2[1], up2[1];
            RunStage<T, 2, 1>(lastRing, z, ctr2, lo1, up1, T::s[2], lo2, up2, s, last);
            lo = lo2[0];
            up = up2[0];
        } else {
            constexpr int n0 = S / T::s[0];
            int lo0[n0], up0[n0];
            RunStage<T, 0, n0>(lds, z, ctr0, rv, rv, T::s[0], lo0, up0, s, last);
            int lo1[1], up1[1];
            RunStage<T, 1, 1>(lastRing, z, ctr1, lo0, up0, T::s[1], lo1, up1, s, last);
            lo = lo1[0];
            up = up1[0];
        }

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
            const float rawScaled = (float)hw_wrap16(rv[j] - k.pedestal) * k.scaler;
            const int bls = hw_round_i16(rawScaled - subBl);
            ol[(size_t)(g * S + j) * LS] = (int16_t)bls;
            const int rr = hw_round_i16(rawScaled);
            const bool maskHp1 = (float)bls < thr;
            const bool m = latH1 && latL && maskHp1;
            latL = (float)bls < thr;
            latH1 = maskHp1;
            // BaselinerStatAccumulator::AddSample on the latent frame
            const float bs = (float)latData;
            if (g == 0 && j < 4) fbuf[j] = bs;
            lagM2 += bbuf[j & 3] * bs;
            bbuf[j & 3] = bs;
            cm0 += 1.0f;
            cm1 += bs;
            cm2 += bs * bs;
            tmin = latData < tmin ? latData : tmin;
            tmax = latData > tmax ? latData : tmax;
            if (m) {
                bm0 = bm0 + 1.0f;
                bm1 = bm1 + bs;
                bm2 = bm2 + bs * bs;
            }
            rawBl += (float)(m ? latRaw : 0);
            latRaw = rr;
            latData = bls;
        }
        const float bgv = (float)hw_round_i16(subBl);
        gm0 += 1.0f;
        gm1 += bgv;
        gm2 += bgv * bgv;
    }

    // Persist carried state.
    __syncthreads();
#pragma unroll 4
    for (int i = 0; i < 4 * T::ldsW; ++i) wl[i * LS + z] = lds[i * LS + z];
#pragma unroll
    for (int f = 0; f < 4 * NST; ++f) {
        st.hold[f][z] = (int16_t)s[f];
        st.last[f][z] = (int16_t)last[f];
    }
    if (z == 0) {
        st.counter[0] = ctr0;
        st.counter[1] = ctr1;
        st.counter[2] = ctr2;
    }
    st.ema_sum[z] = emaSum;
    st.ema_weight[z] = emaWeight;
    st.sigma_ema[z] = sigmaEma;
    st.lat_data[z] = (int16_t)latData;
    st.lat_raw[z] = (int16_t)latRaw;
    st.lat_l[z] = latL ? 1 : 0;
    st.lat_h1[z] = latH1 ? 1 : 0;

    hwpa_baseliner_stats& o = stats[lane];
    o.full.basic.offset[z] = 0.f;
    o.full.basic.m0[z] = cm0;
    o.full.basic.m1[z] = cm1;
    o.full.basic.m2[z] = cm2;
    o.full.lag_m2[z] = lagM2;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        o.full.fbuf[q][z] = fbuf[q];
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
    o.background.m0[z] = gm0;
    o.background.m1[z] = gm1;
    o.background.m2[z] = gm2;
}

__global__ void BaselineInitKernel(BlLaneState* state, int16_t* win, int winPerLane)
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
    for (int f = 0; f < 4 * HWPA_MAX_FILTER_STAGES; ++f) {
        st.hold[f][z] = 0;
        st.last[f][z] = 0;
    }
    st.counter[z] = 0;
    int16_t* wl = win + (size_t)lane * winPerLane * LS;
    for (int i = 0; i < winPerLane; ++i) wl[i * LS + z] = 0;
}

using TMSL = Table<3, 1, 2, 8, 11, 17, 61>;
using TMSM = Table<3, 1, 2, 8, 9, 17, 31>;
using TMSS = Table<3, 1, 2, 8, 7, 17, 17>;
using TTSL = Table<2, 2, 8, 1, 11, 61, 1>;
using TTSM = Table<2, 2, 8, 1, 9, 31, 1>;
using TTSS = Table<2, 2, 8, 1, 7, 17, 1>;

template <class T>
static int32_t Launch(BaselinerDev* b, const int16_t* raw, uint32_t frames, int16_t* out,
                      hwpa_baseliner_stats* stats, hipStream_t s)
{
    if (frames % T::S != 0) return Fail(HWPA_ERR_INVALID_ARG, "frames must be a multiple of the aggregate stride");
    hipLaunchKernelGGL(BaselineKernel<T>, dim3(b->lanes), dim3(LS), 0, s, raw, (int)frames, out, stats,
                       b->state, b->win, b->k);
    return CheckLaunch("BaselineKernel");
}

uint32_t BaselinerStartupLatency(const hwpa_config& cfg)
{
    // MultiScaleBaseliner::StartupLatency (MultiScaleBaseliner.h:56-99)
    const hwpa_filter_params fp = FilterParams(cfg.filter_type);
    const float meanAlpha = hw_powf(0.5f, 1.0f / cfg.mean_ema_scale_strides);
    const float sigmaAlpha = hw_exp2f(-1.0f / cfg.sigma_ema_scale_strides);
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
    b->k.cMeanBias = b->fp.mean_bias * hw_exp2f(cfg.mean_bias_adjust);
    b->k.cSigmaBias = b->fp.sigma_bias * hw_exp2f(cfg.sigma_bias_adjust);
    b->k.meanEmaAlpha = hw_powf(0.5f, 1.0f / cfg.mean_ema_scale_strides);
    b->k.sigmaEmaAlpha = hw_exp2f(-1.0f / cfg.sigma_ema_scale_strides);
    b->k.jumpTol = (cfg.jump_suppression > 0.0f) ? 1.0f / cfg.jump_suppression : INFINITY;
    b->k.scaler = cfg.movie.photoelectron_sensitivity;
    b->k.pedestal = cfg.movie.pedestal;
    size_t w = 0;
    for (int i = 0; i < b->fp.num_stages; ++i) w += 4 * (size_t)b->fp.widths[i];
    b->winPerLane = w;
    int32_t rc = DeviceAlloc((void**)&b->state, sizeof(BlLaneState) * lanes, "baseliner state");
    if (rc == HWPA_OK) rc = DeviceAlloc((void**)&b->win, w * LS * sizeof(int16_t) * lanes, "baseliner windows");
    if (rc != HWPA_OK) {
        BaselinerDestroy(b);
        return rc;
    }
    hipLaunchKernelGGL(BaselineInitKernel, dim3(lanes), dim3(LS), 0, nullptr, b->state, b->win, (int)w);
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

int32_t BaselinerRun(BaselinerDev* b, const int16_t* raw, uint32_t frames, int16_t* out,
                     hwpa_baseliner_stats* stats, hipStream_t s)
{
    if (!b || !raw || !out || !stats || frames == 0) return Fail(HWPA_ERR_INVALID_ARG, "baseliner: bad argument");
    switch (b->fp.num_stages == 3 ? b->fp.widths[0] * 100 + b->fp.widths[2] : 1000 + b->fp.widths[0] * 100 + b->fp.widths[1]) {
    case 1161: return Launch<TMSL>(b, raw, frames, out, stats, s);
    case 931: return Launch<TMSM>(b, raw, frames, out, stats, s);
    case 717: return Launch<TMSS>(b, raw, frames, out, stats, s);
    case 2161: return Launch<TTSL>(b, raw, frames, out, stats, s);
    case 1931: return Launch<TTSM>(b, raw, frames, out, stats, s);
    case 1717: return Launch<TTSS>(b, raw, frames, out, stats, s);
    default: return Fail(HWPA_ERR_UNSUPPORTED, "baseliner: unsupported filter table");
    }
}

}  // namespace hwpa
