// histogram.hip -- per-ZMW 300-bin trace histograms with edge-frame
// scrubbing (TraceHistogramAccumHost, traceAnalysis/TraceHistogramAccumHost.cpp:92-169;
// UHistogramSimd binning, dataTypes/UHistogramSimd.{h,cpp}; EdgeFrameClassifier,
// traceAnalysis/EdgeFrameClassifier.h).
//
// Three forms, all bit-exact, chosen per shape by the autotuner:
//  * HistAddKernel: one wave64 per lane, one thread per ZMW.  The counters of bins
//    [0, NL) -- where baseline and pulses of ordinary signal levels land; NL is
//    128, 160 or 192 (16, 20 or 24 KB of LDS per lane) -- are staged in LDS for
//    the chunk; higher bins are counted with global atomics in place.  Each thread owns one
//    ZMW column, and the [bin pair][zmw] layout puts the 64 lanes of every
//    increment on distinct LDS banks whatever the bins.  Frames are read in
//    batches with the next batch in flight while one is binned.
//  * HistAddWavesKernel: 2, 4 or 8 waves per lane share its LDS counters, each wave
//    taking a contiguous share of the chunk's frames.
//  * HistAddFullKernel: every one of the 300 bins in LDS, the bin found without a
//    divide (hw_shortbin) for all but the samples next to a bin edge, frames from a
//    ring of batches refilled in place ("all 300 bins in LDS" below).

#include "internal.h"

namespace hwpa {

// Edge classifier state per lane: the previous frame, a 2-deep history of
// baseline masks (circular, lane-uniform indices) -- never reset by
// histogram resets.
struct EdgeLane {
    int16_t prev[LS];
    uint8_t hist[2][LS];
    int32_t i0;
    int32_t i1;
    int32_t needWrap;  // set by the fast pass for lanes left to the 16-bit wrapping pass
    int32_t fresh;     // the counters are known to be zero (a reset that left them unwritten)
    int32_t pad[28];
};

size_t EdgeStateBytes() { return sizeof(EdgeLane); }

__global__ __launch_bounds__(64) void EdgeInitKernel(EdgeLane* e)
{
    EdgeLane& s = e[blockIdx.x];
    const int z = threadIdx.x;
    s.prev[z] = 0;
    s.hist[0][z] = 0;
    s.hist[1][z] = 0;
    if (z == 0) {
        s.i0 = 0;
        s.i1 = -1;
        s.needWrap = 0;
        s.fresh = 0;
    }
}

int32_t EdgeStateInit(void* edge, uint32_t lanes, hipStream_t s)
{
    if (!edge || lanes == 0) return Fail(HWPA_ERR_INVALID_ARG, "edge_state_init: bad argument");
    hipLaunchKernelGGL(EdgeInitKernel, dim3(lanes), dim3(LS), 0, s, (EdgeLane*)edge);
    return CheckLaunch("EdgeInitKernel");
}

struct HistConsts {
    float binSizeCoeff, binSizeMin, fallbackSigma, minFrameCount;
};

// TraceHistogramAccumHost::ResetImpl(const BaselinerMetrics&) + UHistogramSimd ctor.
// ZERO writes the 300 zero counters; otherwise the lane is marked fresh and the next
// add starts its counters from zero without reading them (the reset's 38 KB of
// zeros per lane are neither written here nor read back there).
template <bool ZERO>
__global__ __launch_bounds__(64) void HistResetKernel(hwpa_lane_hist* hist, EdgeLane* edge,
                                                      const hwpa_baseliner_stats* stats, HistConsts k)
{
    hwpa_lane_hist& h = hist[blockIdx.x];
    const hwpa_baseliner_stats& s = stats[blockIdx.x];
    const int z = threadIdx.x;
    const float m0 = s.baseline.m0[z], m1 = s.baseline.m1[z], m2 = s.baseline.m2[z], off = s.baseline.offset[z];
    const bool suff = m0 >= k.minFrameCount;
    float var = m1 * m1 / m0;
    var = (m2 - var) / (m0 - 1.0f);
    var = hw_maxf(var, 0.0f);
    var = (m0 > 1.0f) ? var : hw_u2f(0x7fc00000u);
    const float mean = suff ? (m1 / m0 + off) : 0.0f;
    const float sigma = suff ? hw_sqrtf(var) : k.fallbackSigma;
    const float binSize = hw_maxf(k.binSizeCoeff * sigma, k.binSizeMin);
    const float lo = mean - 4.0f * sigma;
    const float up = lo + (float)NB * binSize;
    h.low_bound[z] = lo;
    h.bin_size[z] = (up - lo) / (float)NB;
    h.outlier_low[z] = 0;
    h.outlier_high[z] = 0;
    if (ZERO) {
        for (int b = 0; b < NB; ++b) h.bin_count[b][z] = 0;
    } else if (z == 0) {
        edge[blockIdx.x].fresh = 1;
    }
}

// Zeroes bins [NL, NB) of a fresh lane before they are counted in place; the
// caller's barrier orders these stores before the increments.
template <int NL, int NT>
__device__ __forceinline__ void ZeroHighBins(hwpa_lane_hist& h, int tid)
{
    if (NL < NB) {
        uint4* rest = reinterpret_cast<uint4*>(&h.bin_count[NL][0]);
        constexpr int NV = (NB - NL) * LS * 2 / 16;
        for (int i = tid; i < NV; i += NT) rest[i] = make_uint4(0u, 0u, 0u, 0u);
    }
}

// Bin of a sample: floor((x - lo) / bs) with the reference's IEEE divide
// (UHistogramSimd::BinIndex, dataTypes/UHistogramSimd.h:199-229).  The
// quotient is first formed as q = (x - lo) * (1 / bs) with 1 / bs rounded
// once; q is within 2 * 2^-24 |y| of y = (x - lo) / bs and the divide's
// result within 2^-24 |y|, so both lie within 3 * 2^-24 |y| < 2^-22 |q| of
// each other.  When the fraction of q is farther than that from 0 and 1 the
// two share their floor (for |q| < 1 the common sign decides).  For
// 2048 <= |q| < 2^30 the sample is an outlier on the same side either way.
// Everything else -- near a bin edge, zero, NaN, infinities, huge quotients
// -- is flagged for the divide (whose float -> int conversion gives
// 0x80000000 out of range, a low outlier, as on the host).
__device__ __forceinline__ int BinFast(float a, float inv, bool& exact)
{
    const float yp = a * inv;
    const float ay = hw_fabsf(yp);
    const bool far = ay >= 2048.0f && ay < 1073741824.0f;
    const float fl = hw_floorf(yp);
    const float d = yp - fl;
    const float tol = __builtin_fmaxf(ay * 2.384185791015625e-07f, 9.5367431640625e-07f);  // 2^-22 |q|, >= 2^-20
    const bool clear = ay < 2048.0f && d >= tol && d <= 1.0f - tol;
    // Branch-free: clamp into [-1, NB] in float, pick the outlier side, convert once.
    float bf = __builtin_fminf(__builtin_fmaxf(fl, -1.0f), (float)NB);
    const float side = (yp > 0.0f) ? (float)NB : -1.0f;
    bf = far ? side : bf;
    exact = !far && !clear;
    return (int)bf;
}

// One batch of N frames of TraceHistogramAccumHost::AddBlock
// (TraceHistogramAccumHost.cpp:128-169): EdgeFrameClassifier::IsEdgeFrame
// then UHistogramSimd::AddDatum per frame.  The classifier's two-entry
// circular array is a shift register here: bl1/bl2 = baseline flags of the
// previous two frames, pushes saturates at 2 (lane-uniform); the previous
// frame counts when the array was full and its two neighbours agree.
// Counters live in LDS as words holding bins 2j (low half) and 2j+1 (high
// half) of one ZMW, so an increment is a single ds_add_u32 that nothing waits
// on (frames that are not counted add 0).  That cannot carry into the
// neighbouring half while every count stays below 65536 -- checked before the
// chunk; otherwise the wave takes the WRAP variant, a 16-bit wrapping
// read-modify-write, the reference's u16 counter semantics.  Bins [0, NL) are
// counted in LDS; bins [NL, NB) with 32-bit global atomics on the lane's
// histogram in HBM, each word holding the 16-bit counters of ZMWs z and z^1
// for one bin (no carry under the same 65536 guard).  The rare divides and
// global atomics are taken once per batch, by the lanes that need them.
template <int N, bool WRAP, int NL>
__device__ __forceinline__ void HistBatch(uint32_t* cnt, hwpa_lane_hist& hg, const int (&fv)[N], int z, float lo,
                                          float bs, float inv, int thr, int& prev, bool& bl1, bool& bl2, int& pushes,
                                          uint32_t& lowOut, uint32_t& highOut)
{
    int bin[N];
    uint32_t countedMask = 0, exactMask = 0;
#pragma clang loop unroll(full)
    for (int u = 0; u < N; ++u) {
        const bool isBl = fv[u] < thr;
        const bool counted = (u >= 2 || pushes + u >= 2) && (isBl == bl2);
        bl2 = bl1;
        bl1 = isBl;
        const int cand = (u == 0) ? prev : fv[u - 1];
        bool ex;
        bin[u] = BinFast((float)cand - lo, inv, ex);
        countedMask |= (counted ? 1u : 0u) << u;
        exactMask |= (ex ? 1u : 0u) << u;
    }
    if (__any(exactMask != 0u)) {
#pragma clang loop unroll(full)
        for (int u = 0; u < N; ++u) {
            if ((exactMask >> u) & 1u) {
                const int cand = (u == 0) ? prev : fv[u - 1];
                bin[u] = hw_floor_i32(((float)cand - lo) / bs);
            }
        }
    }
    uint32_t globMask = 0, nLow = 0, nHigh = 0;
#pragma clang loop unroll(full)
    for (int u = 0; u < N; ++u) {
        const bool counted = (countedMask >> u) & 1u;
        const int b = bin[u];
        const bool low = b < 0, high = b >= NB;
        nLow += (counted && low) ? 1u : 0u;
        nHigh += (counted && high) ? 1u : 0u;
        const bool in = counted && !low && !high;
        const int bc = low ? 0 : (high ? NB - 1 : b);
        if (NL < NB) globMask |= ((in && bc >= NL) ? 1u : 0u) << u;
        const int bl = (NL < NB && bc >= NL) ? NL - 1 : bc;
        uint32_t* w = &cnt[(bl >> 1) * LS + z];
        const uint32_t sh = (uint32_t)(bl & 1) * 16u;
        const bool inl = in && (NL == NB || bc < NL);
        if (!WRAP) {
            atomicAdd(w, inl ? (1u << sh) : 0u);
        } else if (inl) {
            const uint32_t v = *w;
            const uint32_t c = ((v >> sh) + 1u) & 0xffffu;
            *w = (v & ~(0xffffu << sh)) | (c << sh);
        }
    }
    if (NL < NB && __any(globMask != 0u)) {
#pragma clang loop unroll(full)
        for (int u = 0; u < N; ++u) {
            if ((globMask >> u) & 1u) {
                uint32_t* gw = reinterpret_cast<uint32_t*>(&hg.bin_count[bin[u]][z & ~1]);
                atomicAdd(gw, 1u << ((uint32_t)(z & 1) * 16u));
            }
        }
    }
    lowOut = (lowOut + nLow) & 0xffffu;
    highOut = (highOut + nHigh) & 0xffffu;
    prev = fv[N - 1];
    pushes = (pushes + N < 2) ? pushes + N : 2;
}

template <int B, bool WRAP, int NL>
__device__ __forceinline__ void HistFrames(uint32_t* cnt, hwpa_lane_hist& hg, const int16_t* tr, int frames, int z,
                                           float lo, float bs, float inv, int thr, int& prev, bool& bl1, bool& bl2,
                                           int& pushes, uint32_t& lowOut, uint32_t& highOut)
{
    const int full = frames / B;
    int fv[B];
    if (full > 0) {
#pragma clang loop unroll(full)
        for (int u = 0; u < B; ++u) fv[u] = tr[(size_t)u * LS];
    }
    for (int g = 0; g < full; ++g) {
        int nv[B];
        const int gn = (g + 1 < full) ? g + 1 : g;
#pragma clang loop unroll(full)
        for (int u = 0; u < B; ++u) nv[u] = tr[(size_t)(gn * B + u) * LS];
        HistBatch<B, WRAP, NL>(cnt, hg, fv, z, lo, bs, inv, thr, prev, bl1, bl2, pushes, lowOut, highOut);
#pragma clang loop unroll(full)
        for (int u = 0; u < B; ++u) fv[u] = nv[u];
    }
    for (int t = full * B; t < frames; ++t) {
        const int f1[1] = {tr[(size_t)t * LS]};
        HistBatch<1, WRAP, NL>(cnt, hg, f1, z, lo, bs, inv, thr, prev, bl1, bl2, pushes, lowOut, highOut);
    }
}

// The word of ZMW pair (z & ~1) that thread z moves between HBM and the
// packed LDS layout: bin 2j for even z, bin 2j+1 for odd z.  One dword per
// thread and pair of bins; the two halves are exchanged with the neighbour.
__device__ __forceinline__ uint32_t UMax(uint32_t a, uint32_t b) { return a > b ? a : b; }

__device__ __forceinline__ uint32_t SwapPair(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, false);  // quad_perm [1,0,3,2]
}

// Two passes so each is compiled on its own: the fast pass runs every lane
// whose counts stay below 65536 and flags the others (leaving them untouched);
// the wrap pass then runs only the flagged lanes.
template <int B, bool WRAP, int NL>
__global__ __launch_bounds__(64) void HistAddKernel(hwpa_lane_hist* hist, EdgeLane* edge,
                                                    const int16_t* __restrict__ traces, int frames,
                                                    const hwpa_lane_model* __restrict__ models)
{
    static_assert(NL % 2 == 0 && NL <= NB, "LDS bins");
    __shared__ uint32_t cnt[(NL / 2) * LS];
    const int lane = blockIdx.x;
    const int z = threadIdx.x;
    hwpa_lane_hist& h = hist[lane];
    EdgeLane& e = edge[lane];
    if (WRAP && e.needWrap == 0) return;
    const bool odd = (z & 1) != 0;
    // LDS bins: thread z reads the dword of bin 2j + (z & 1) holding ZMWs
    // z & ~1 and (z & ~1) + 1, 16 bin pairs per batch so the loads overlap.
    const uint32_t* rows = reinterpret_cast<const uint32_t*>(&h.bin_count[0][0]);
    const int col = (z >> 1) + (odd ? LS / 2 : 0);   // dword index within a bin pair (2 rows = 64 dwords)
    uint32_t mx = 0;
    constexpr int NJ = NL / 2;
    const bool fresh = !WRAP && e.fresh != 0;
    if (fresh) {
        for (int j = 0; j < NJ; ++j) cnt[j * LS + z] = 0u;
        ZeroHighBins<NL, LS>(h, z);
        __threadfence_block();
    }
#pragma unroll 1
    for (int j0 = 0; j0 < (fresh ? 0 : NJ); j0 += 16) {
        uint32_t c[16];
#pragma clang loop unroll(full)
        for (int k = 0; k < 16; ++k) {
            const int j = (j0 + k < NJ) ? j0 + k : NJ - 1;
            c[k] = rows[j * LS + col];
        }
#pragma clang loop unroll(full)
        for (int k = 0; k < 16; ++k) {
            const uint32_t v = c[k], n = SwapPair(v);
            mx = UMax(mx, UMax(v & 0xffffu, v >> 16));
            const uint32_t packed = odd ? ((n >> 16) | (v & 0xffff0000u)) : ((v & 0xffffu) | (n << 16));
            const int j = (j0 + k < NJ) ? j0 + k : NJ - 1;
            cnt[j * LS + z] = packed;
        }
    }
    // Bins [NL, NB): only their maximum matters; read as 16-byte vectors.
    if (NL < NB && !fresh) {
        const uint4* rest = reinterpret_cast<const uint4*>(&h.bin_count[NL][0]);
        constexpr int NV = (NB - NL) * LS * 2 / 16;
#pragma unroll 1
        for (int i0 = 0; i0 < NV; i0 += 8 * LS) {
            uint4 c[8];
#pragma clang loop unroll(full)
            for (int k = 0; k < 8; ++k) {
                const int i = i0 + k * LS + z;
                c[k] = rest[i < NV ? i : NV - 1];
            }
#pragma clang loop unroll(full)
            for (int k = 0; k < 8; ++k) {
                const uint32_t w[4] = {c[k].x, c[k].y, c[k].z, c[k].w};
#pragma clang loop unroll(full)
                for (int q = 0; q < 4; ++q)
                    mx = UMax(mx, UMax(w[q] & 0xffffu, w[q] >> 16));
            }
        }
    }
    const bool fast = __all(mx + (uint32_t)frames <= 65535u);
    if (!WRAP && !fast) {
        if (z == 0) e.needWrap = 1;
        return;
    }
    uint32_t lowOut = h.outlier_low[z], highOut = h.outlier_high[z];
    const float lo = h.low_bound[z];
    const float bs = h.bin_size[z];
    const float inv = 1.0f / bs;

    // Edge threshold from the current model (TraceHistogramAccumHost.cpp:156-157)
    const hwpa_lane_model& m = models[lane];
    const float bvar = hw_h2f(m.baseline.vars[z]);
    const float bmean = hw_h2f(m.baseline.means[z]);
    const int thr = hw_wrap16(hw_round_i32(2.0f * hw_sqrtf(bvar) + bmean));

    // CircularArray state (h[2], head i0, tail i1; i1 < 0 when empty) -> shift register.
    int prev = e.prev[z];
    const int i0 = e.i0, i1 = e.i1;
    int pushes = (i1 < 0) ? 0 : ((2 + i1 - i0) & 1) + 1;
    bool bl1 = (i1 >= 0) && e.hist[i1 & 1][z] != 0;
    bool bl2 = (pushes == 2) && e.hist[i0 & 1][z] != 0;

    const int16_t* tr = traces + (size_t)lane * frames * LS + z;
    HistFrames<B, WRAP, NL>(cnt, h, tr, frames, z, lo, bs, inv, thr, prev, bl1, bl2, pushes, lowOut, highOut);

    // Back to HBM, one dword per thread and bin pair (the neighbour's halves
    // swapped in).  Bins >= NL were counted in place.
    uint32_t* wrows = reinterpret_cast<uint32_t*>(&h.bin_count[0][0]);
#pragma unroll 4
    for (int j = 0; j < NJ; ++j) {
        const uint32_t v = cnt[j * LS + z], n = SwapPair(v);
        wrows[j * LS + col] = odd ? ((n >> 16) | (v & 0xffff0000u)) : ((v & 0xffffu) | (n << 16));
    }
    h.outlier_low[z] = (uint16_t)lowOut;
    h.outlier_high[z] = (uint16_t)highOut;
    // Back to the circular-array form: oldest flag at the head.
    e.prev[z] = (int16_t)prev;
    e.hist[0][z] = (pushes == 2 ? bl2 : bl1) ? 1 : 0;
    e.hist[1][z] = bl1 ? 1 : 0;
    if (z == 0) {
        e.i0 = 0;
        e.i1 = (pushes == 0) ? -1 : (pushes == 1 ? 0 : 1);
        e.needWrap = 0;
        e.fresh = 0;
    }
}

// W waves per lane, sharing the lane's LDS counters: wave w bins frames [w len, (w + 1) len)
// of the chunk (len = frames / W rounded up to whole batches), so W times as many trace
// rows are in flight per lane.  The edge classifier decides a frame from the frame and
// its two predecessors only, so every wave but the first starts from the two trace frames
// before its range (a full two-entry history); the first continues the carried state and
// the wave holding the last frame stores it.  LDS increments from different waves are
// atomic; the outlier counts are partial sums per wave, added mod 2^16 at the end.  Lanes
// whose counts could pass 65535 are left to the one-wave wrapping pass, as above.
template <int B, int NL, int W>
__global__ __launch_bounds__(64 * W) void HistAddWavesKernel(hwpa_lane_hist* hist, EdgeLane* edge,
                                                             const int16_t* __restrict__ traces, int frames,
                                                             const hwpa_lane_model* __restrict__ models)
{
    static_assert(NL % 2 == 0 && NL <= NB && W >= 2, "LDS bins, waves");
    __shared__ uint32_t cnt[(NL / 2) * LS];
    __shared__ uint32_t outl[2 * LS];  // this chunk's low / high outlier counts
    const int lane = blockIdx.x;
    const int z = threadIdx.x % LS;
    const int w = threadIdx.x / LS;
    hwpa_lane_hist& h = hist[lane];
    EdgeLane& e = edge[lane];
    for (int i = threadIdx.x; i < 2 * LS; i += W * LS) outl[i] = 0;
    const bool odd = (z & 1) != 0;
    const uint32_t* rows = reinterpret_cast<const uint32_t*>(&h.bin_count[0][0]);
    const int col = (z >> 1) + (odd ? LS / 2 : 0);
    uint32_t mx = 0;
    constexpr int NJ = NL / 2;
    const bool fresh = e.fresh != 0;
    if (fresh) {
        for (int j = w; j < NJ; j += W) cnt[j * LS + z] = 0u;
        ZeroHighBins<NL, W * LS>(h, threadIdx.x);
        __threadfence_block();
    }
#pragma unroll 1
    for (int j0 = 16 * w; j0 < (fresh ? 0 : NJ); j0 += 16 * W) {
        uint32_t c[16];
#pragma clang loop unroll(full)
        for (int k = 0; k < 16; ++k) {
            const int j = (j0 + k < NJ) ? j0 + k : NJ - 1;
            c[k] = rows[j * LS + col];
        }
#pragma clang loop unroll(full)
        for (int k = 0; k < 16; ++k) {
            const uint32_t v = c[k], n = SwapPair(v);
            mx = UMax(mx, UMax(v & 0xffffu, v >> 16));
            const uint32_t packed = odd ? ((n >> 16) | (v & 0xffff0000u)) : ((v & 0xffffu) | (n << 16));
            const int j = (j0 + k < NJ) ? j0 + k : NJ - 1;
            cnt[j * LS + z] = packed;   // a clamped duplicate writes the same value
        }
    }
    if (NL < NB && !fresh) {
        const uint4* rest = reinterpret_cast<const uint4*>(&h.bin_count[NL][0]);
        constexpr int NV = (NB - NL) * LS * 2 / 16;
#pragma unroll 1
        for (int i0 = 8 * LS * w; i0 < NV; i0 += 8 * LS * W) {
            uint4 c[8];
#pragma clang loop unroll(full)
            for (int k = 0; k < 8; ++k) {
                const int i = i0 + k * LS + z;
                c[k] = rest[i < NV ? i : NV - 1];
            }
#pragma clang loop unroll(full)
            for (int k = 0; k < 8; ++k) {
                const uint32_t wd[4] = {c[k].x, c[k].y, c[k].z, c[k].w};
#pragma clang loop unroll(full)
                for (int q = 0; q < 4; ++q)
                    mx = UMax(mx, UMax(wd[q] & 0xffffu, wd[q] >> 16));
            }
        }
    }
    // Barrier for the counters and the outlier sums; any wave with a count that could wrap
    // sends the whole lane to the wrapping pass.
    if (__syncthreads_or(mx + (uint32_t)frames > 65535u ? 1 : 0)) {
        if (threadIdx.x == 0) e.needWrap = 1;
        return;
    }
    const float lo = h.low_bound[z];
    const float bs = h.bin_size[z];
    const float inv = 1.0f / bs;
    const hwpa_lane_model& m = models[lane];
    const float bvar = hw_h2f(m.baseline.vars[z]);
    const float bmean = hw_h2f(m.baseline.means[z]);
    const int thr = hw_wrap16(hw_round_i32(2.0f * hw_sqrtf(bvar) + bmean));

    const int len = ((frames + W - 1) / W + B - 1) / B * B;   // >= B >= 2: the launcher needs frames >= W B
    const int f0 = w * len;
    const int n = (frames - f0 < len) ? frames - f0 : len;   // <= 0 for waves past the end
    const int16_t* tr = traces + (size_t)lane * frames * LS + z;
    int prev, pushes;
    bool bl1, bl2;
    if (w == 0) {
        prev = e.prev[z];
        const int i0 = e.i0, i1 = e.i1;
        pushes = (i1 < 0) ? 0 : ((2 + i1 - i0) & 1) + 1;
        bl1 = (i1 >= 0) && e.hist[i1 & 1][z] != 0;
        bl2 = (pushes == 2) && e.hist[i0 & 1][z] != 0;
    } else {
        const int fs = (n > 0) ? f0 : 2;   // waves past the end read in-range frames they do not use
        prev = tr[(size_t)(fs - 1) * LS];
        bl1 = prev < thr;
        bl2 = tr[(size_t)(fs - 2) * LS] < thr;
        pushes = 2;
    }
    uint32_t lowP = 0, highP = 0;
    if (n > 0)
        HistFrames<B, false, NL>(cnt, h, tr + (size_t)f0 * LS, n, z, lo, bs, inv, thr, prev, bl1, bl2, pushes, lowP,
                                 highP);
    atomicAdd(&outl[z], lowP);
    atomicAdd(&outl[LS + z], highP);
    __syncthreads();

    uint32_t* wrows = reinterpret_cast<uint32_t*>(&h.bin_count[0][0]);
#pragma unroll 4
    for (int j = w; j < NJ; j += W) {
        const uint32_t v = cnt[j * LS + z], nb = SwapPair(v);
        wrows[j * LS + col] = odd ? ((nb >> 16) | (v & 0xffff0000u)) : ((v & 0xffffu) | (nb << 16));
    }
    if (w == 0) {
        h.outlier_low[z] = (uint16_t)((uint32_t)h.outlier_low[z] + outl[z]);
        h.outlier_high[z] = (uint16_t)((uint32_t)h.outlier_high[z] + outl[LS + z]);
    }
    if (w == (frames - 1) / len) {
        e.prev[z] = (int16_t)prev;
        e.hist[0][z] = (pushes == 2 ? bl2 : bl1) ? 1 : 0;
        e.hist[1][z] = bl1 ? 1 : 0;
        if (z == 0) {
            e.i0 = 0;
            e.i1 = (pushes == 0) ? -1 : (pushes == 1 ? 0 : 1);
            e.needWrap = 0;
        }
    }
    if (threadIdx.x == 0) e.fresh = 0;
}

// ---- all 300 bins in LDS ---------------------------------------------------------------
// The lane's counters sit in LDS in the HBM layout ([bin][zmw] u16, ZMWs 2i and 2i+1 of a bin
// sharing a 32-bit word), 38,400 bytes, so they move between LDS and HBM as 16-byte vectors
// and a sample of ZMW z in bin b is one ds_add_u32 of 1 << (16 (z & 1)) to the word of
// (b, z & ~1).  No carry crosses into the neighbour's half while every count stays below
// 65536 (checked before the chunk, as above).
//
// Bin of a sample without a divide: hw_shortbin (include/hwpa_math.h), which decides every
// sample except those within a rounding band of a bin edge, outliers and unusable bounds;
// such samples take the exact path (BinFast / the divide and the outlier counts), which a
// batch runs only for the samples that need it after the whole batch has been counted.

template <int N, class V>
__device__ __forceinline__ void HistBatchFull(char* cntB, uint32_t zoff, uint32_t inc, const V (&fv)[N],
                                              const hw_shortbin_z& zb, int thr, int& prev, bool& bl1, bool& bl2,
                                              int& pushes, uint32_t& lowOut, uint32_t& highOut)
{
    const bool s1 = bl1, s2 = bl2;
    const int sp = pushes, sprev = prev;
    bool slow = false;
#pragma clang loop unroll(full)
    for (int u = 0; u < N; ++u) {
        const bool isBl = (int)fv[u] < thr;
        const bool counted = (u >= 2 || pushes + u >= 2) && (isBl == bl2);
        bl2 = bl1;
        bl1 = isBl;
        const int cand = (u == 0) ? prev : fv[u - 1];
        int b;
        const bool ok = hw_shortbin(cand, zb, NB, &b);
        slow = slow || (counted && !ok);
        const uint32_t bc = ok ? (uint32_t)b : 0u;   // a sample off the short path adds 0 to bin 0
        atomicAdd(reinterpret_cast<uint32_t*>(cntB + bc * (uint32_t)(LS * 2) + zoff), (counted && ok) ? inc : 0u);
    }
    if (__any(slow)) {
        bool b1 = s1, b2 = s2;
        uint32_t nLow = 0, nHigh = 0;
#pragma clang loop unroll(full)
        for (int u = 0; u < N; ++u) {
            const bool isBl = (int)fv[u] < thr;
            const bool counted = (u >= 2 || sp + u >= 2) && (isBl == b2);
            b2 = b1;
            b1 = isBl;
            const int cand = (u == 0) ? sprev : fv[u - 1];
            int sb;
            if (counted && !hw_shortbin(cand, zb, NB, &sb)) {
                bool ex;
                int bin = BinFast((float)cand - zb.lo, zb.inv, ex);
                if (ex) bin = hw_floor_i32(((float)cand - zb.lo) / zb.bs);
                if (bin < 0) {
                    ++nLow;
                } else if (bin >= NB) {
                    ++nHigh;
                } else {
                    atomicAdd(reinterpret_cast<uint32_t*>(cntB + (uint32_t)bin * (uint32_t)(LS * 2) + zoff), inc);
                }
            }
        }
        lowOut = (lowOut + nLow) & 0xffffu;
        highOut = (highOut + nHigh) & 0xffffu;
    }
    prev = fv[N - 1];
    pushes = (pushes + N < 2) ? pushes + N : 2;
}

// Frames in batches of B with PD batches in flight.  Slot j of the register ring holds batch
// g0 + j; the loop body always covers PD whole batches, binning each slot and refilling it
// in place with batch g0 + j + PD, so no register is copied while its load is in flight
// (a copy would make the wave wait for every outstanding load) and each batch's loads have
// PD - 1 batches of work to land.  The last full % PD batches, already loaded, follow.
// Prime() issues the first PD batches; the kernel calls it before its other set-up loads so
// their latencies overlap.
template <int B, int PD>
struct FrameRing {
    const int16_t* tr;
    int frames, full;
    // Raw samples: the loads of the set-up and of the loop are then the same instruction, so
    // the ring keeps its registers across the loop (no copy that would wait for the loads).
    int16_t fv[PD][B];

    __device__ __forceinline__ void Load(int j, int g)
    {
        g = (g < full) ? g : full - 1;   // past the end: re-read the last batch (never binned)
#pragma clang loop unroll(full)
        for (int u = 0; u < B; ++u) fv[j][u] = __builtin_nontemporal_load(&(tr + (size_t)g * B * LS)[u * LS]);
    }
    __device__ __forceinline__ void Prime(const int16_t* t, int n)
    {
        tr = t;
        frames = n;
        full = (n > 0) ? n / B : 0;
        if (full > 0) {
#pragma clang loop unroll(full)
            for (int j = 0; j < PD; ++j) Load(j, j);
        }
    }
    __device__ __forceinline__ void Run(char* cntB, uint32_t zoff, uint32_t inc, const hw_shortbin_z& zb, int thr, int& prev,
                                        bool& bl1, bool& bl2, int& pushes, uint32_t& lowOut, uint32_t& highOut)
    {
        if (full > 0) {
            int g0 = 0;
            for (; g0 + PD <= full; g0 += PD) {
#pragma clang loop unroll(full)
                for (int j = 0; j < PD; ++j) {
                    HistBatchFull<B>(cntB, zoff, inc, fv[j], zb, thr, prev, bl1, bl2, pushes, lowOut, highOut);
                    Load(j, g0 + j + PD);
                }
            }
#pragma clang loop unroll(full)
            for (int j = 0; j + 1 < PD; ++j) {
                if (g0 + j < full)
                    HistBatchFull<B>(cntB, zoff, inc, fv[j], zb, thr, prev, bl1, bl2, pushes, lowOut, highOut);
            }
        }
        for (int t = full * B; t < frames; ++t) {
            const int f1[1] = {tr[(size_t)t * LS]};
            HistBatchFull<1>(cntB, zoff, inc, f1, zb, thr, prev, bl1, bl2, pushes, lowOut, highOut);
        }
    }
};

typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t PkMaxU16(uint32_t a, uint32_t b)
{
    return __builtin_bit_cast(uint32_t,
                              __builtin_elementwise_max(__builtin_bit_cast(u16x2, a), __builtin_bit_cast(u16x2, b)));
}

// W waves per lane, as HistAddWavesKernel, with every bin in LDS; PD batches of B frames in
// flight per wave.
template <int B, int W, int PD>
__global__ __launch_bounds__(64 * W) void HistAddFullKernel(hwpa_lane_hist* hist, EdgeLane* edge,
                                                            const int16_t* __restrict__ traces, int frames,
                                                            const hwpa_lane_model* __restrict__ models)
{
    static_assert(W >= 2, "waves");
    constexpr int NV = NB * LS * 2 / 16;   // 16-byte vectors of counters
    __shared__ uint4 cnt[NV];
    __shared__ uint32_t outl[2 * LS];
    const int lane = blockIdx.x;
    const int z = threadIdx.x % LS;
    const int w = threadIdx.x / LS;
    hwpa_lane_hist& h = hist[lane];
    EdgeLane& e = edge[lane];

    // Every load the set-up needs is issued here, the wave's first PD trace batches first.
    const int len = ((frames + W - 1) / W + B - 1) / B * B;   // >= B >= 2: the launcher needs frames >= W B
    const int f0 = w * len;
    const int n = (frames - f0 < len) ? frames - f0 : len;   // <= 0 for waves past the end
    const int16_t* tr = traces + (size_t)lane * frames * LS + z;
    FrameRing<B, PD> ring;
    ring.Prime(tr + (size_t)(n > 0 ? f0 : 0) * LS, n);
    const int fs = (n > 0) ? f0 : 2;   // waves past the end read in-range frames they do not use
    const int before1 = (w > 0) ? (int)tr[(size_t)(fs - 1) * LS] : 0;
    const int before2 = (w > 0) ? (int)tr[(size_t)(fs - 2) * LS] : 0;
    const int fresh = e.fresh;
    const float lo = h.low_bound[z], bs = h.bin_size[z];
    const hwpa_lane_model& m = models[lane];
    const float bvar = hw_h2f(m.baseline.vars[z]);
    const float bmean = hw_h2f(m.baseline.means[z]);

    for (int i = threadIdx.x; i < 2 * LS; i += W * LS) outl[i] = 0;
    uint4* hv = reinterpret_cast<uint4*>(&h.bin_count[0][0]);
    uint32_t mx = 0;   // two 16-bit maxima
    if (fresh != 0) {
        for (int i = threadIdx.x; i < NV; i += W * LS) cnt[i] = make_uint4(0u, 0u, 0u, 0u);
    } else {
        constexpr int K = 4;
#pragma unroll 1
        for (int i0 = threadIdx.x; i0 < NV; i0 += K * W * LS) {
            uint4 c[K];
#pragma clang loop unroll(full)
            for (int k = 0; k < K; ++k) {
                const int i = i0 + k * W * LS;
                c[k] = hv[i < NV ? i : NV - 1];
            }
#pragma clang loop unroll(full)
            for (int k = 0; k < K; ++k) {
                const int i = i0 + k * W * LS;
                if (i < NV) cnt[i] = c[k];
                mx = PkMaxU16(mx, PkMaxU16(PkMaxU16(c[k].x, c[k].y), PkMaxU16(c[k].z, c[k].w)));
            }
        }
    }
    const uint32_t mxc = UMax(mx & 0xffffu, mx >> 16);
    if (__syncthreads_or(mxc + (uint32_t)frames > 65535u ? 1 : 0)) {
        if (threadIdx.x == 0) e.needWrap = 1;   // counters untouched in HBM; the wrapping pass takes the lane
        __builtin_amdgcn_s_waitcnt(0x0F70);      // the primed trace loads land before the wave ends
        return;
    }
    const hw_shortbin_z zb = hw_shortbin_make(lo, bs);
    const int thr = hw_wrap16(hw_round_i32(2.0f * hw_sqrtf(bvar) + bmean));
    const uint32_t zoff = (uint32_t)(z & ~1) * 2u;
    const uint32_t inc = 1u << (16u * (uint32_t)(z & 1));
    char* cntB = reinterpret_cast<char*>(cnt);

    int prev, pushes;
    bool bl1, bl2;
    if (w == 0) {
        prev = e.prev[z];
        const int i0 = e.i0, i1 = e.i1;
        pushes = (i1 < 0) ? 0 : ((2 + i1 - i0) & 1) + 1;
        bl1 = (i1 >= 0) && e.hist[i1 & 1][z] != 0;
        bl2 = (pushes == 2) && e.hist[i0 & 1][z] != 0;
    } else {
        prev = before1;
        bl1 = before1 < thr;
        bl2 = before2 < thr;
        pushes = 2;
    }
    uint32_t lowP = 0, highP = 0;
    ring.Run(cntB, zoff, inc, zb, thr, prev, bl1, bl2, pushes, lowP, highP);
    atomicAdd(&outl[z], lowP);
    atomicAdd(&outl[LS + z], highP);
    __syncthreads();

    typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
    for (int i = threadIdx.x; i < NV; i += W * LS) {
        const uint4 c = cnt[i];
        const u32x4 v = {c.x, c.y, c.z, c.w};
        __builtin_nontemporal_store(v, reinterpret_cast<u32x4*>(&hv[i]));
    }
    if (w == 0) {
        h.outlier_low[z] = (uint16_t)((uint32_t)h.outlier_low[z] + outl[z]);
        h.outlier_high[z] = (uint16_t)((uint32_t)h.outlier_high[z] + outl[LS + z]);
    }
    if (w == (frames - 1) / len) {
        e.prev[z] = (int16_t)prev;
        e.hist[0][z] = (pushes == 2 ? bl2 : bl1) ? 1 : 0;
        e.hist[1][z] = bl1 ? 1 : 0;
        if (z == 0) {
            e.i0 = 0;
            e.i1 = (pushes == 0) ? -1 : (pushes == 1 ? 0 : 1);
            e.needWrap = 0;
        }
    }
    if (threadIdx.x == 0) e.fresh = 0;
}

static HistConsts MakeHistConsts(const hwpa_config& cfg)
{
    HistConsts k;
    k.binSizeCoeff = cfg.bin_size_coeff;
    k.binSizeMin = cfg.bin_size_low_bound_coeff * std::max(1.0f, cfg.movie.photoelectron_sensitivity);
    k.fallbackSigma = cfg.fallback_baseline_sigma;
    k.minFrameCount = (float)cfg.baseline_stat_min_frame_count;
    return k;
}

// Reset that defers the zero counters to the lane's next add (which must come
// before anything reads the counters, with at most 65535 frames).
int32_t HistResetDeferred(const hwpa_config& cfg, hwpa_lane_hist* hist, void* edge, const hwpa_baseliner_stats* stats,
                          uint32_t lanes, hipStream_t s)
{
    if (!hist || !edge || !stats || lanes == 0) return Fail(HWPA_ERR_INVALID_ARG, "hist_reset: bad argument");
    hipLaunchKernelGGL((HistResetKernel<false>), dim3(lanes), dim3(LS), 0, s, hist, (EdgeLane*)edge, stats,
                       MakeHistConsts(cfg));
    return CheckLaunch("HistResetKernel");
}

int32_t HistReset(const hwpa_config& cfg, hwpa_lane_hist* hist, const hwpa_baseliner_stats* stats, uint32_t lanes,
                  hipStream_t s)
{
    if (!hist || !stats || lanes == 0) return Fail(HWPA_ERR_INVALID_ARG, "hist_reset: bad argument");
    hipLaunchKernelGGL((HistResetKernel<true>), dim3(lanes), dim3(LS), 0, s, hist, nullptr, stats,
                       MakeHistConsts(cfg));
    return CheckLaunch("HistResetKernel");
}

// The 16-bit wrapping pass over the lanes a fast pass left to it (edge state needWrap).
static int32_t HistWrapLaunch(hwpa_lane_hist* hist, EdgeLane* edge, const int16_t* traces, uint32_t frames,
                              const hwpa_lane_model* models, uint32_t lanes, hipStream_t s)
{
    hipLaunchKernelGGL((HistAddKernel<16, true, NB>), dim3(lanes), dim3(LS), 0, s, hist, edge, traces, (int)frames,
                       models);
    return CheckLaunch("HistAddKernel");
}

template <int B, int NL>
static int32_t HistAddLaunch(hwpa_lane_hist* hist, EdgeLane* edge, const int16_t* traces, uint32_t frames,
                             const hwpa_lane_model* models, uint32_t lanes, hipStream_t s, bool wrap)
{
    hipLaunchKernelGGL((HistAddKernel<B, false, NL>), dim3(lanes), dim3(LS), 0, s, hist, edge, traces, (int)frames,
                       models);
    const int32_t rc = CheckLaunch("HistAddKernel");
    if (rc != HWPA_OK || !wrap) return rc;
    return HistWrapLaunch(hist, edge, traces, frames, models, lanes, s);
}

template <int B, int NL, int W>
static int32_t HistAddWavesLaunch(hwpa_lane_hist* hist, EdgeLane* edge, const int16_t* traces, uint32_t frames,
                                  const hwpa_lane_model* models, uint32_t lanes, hipStream_t s, bool wrap)
{
    if (frames < (uint32_t)(W * B)) return HistAddLaunch<B, NL>(hist, edge, traces, frames, models, lanes, s, wrap);
    hipLaunchKernelGGL((HistAddWavesKernel<B, NL, W>), dim3(lanes), dim3(W * LS), 0, s, hist, edge, traces,
                       (int)frames, models);
    const int32_t rc = CheckLaunch("HistAddWavesKernel");
    if (rc != HWPA_OK || !wrap) return rc;
    return HistWrapLaunch(hist, edge, traces, frames, models, lanes, s);
}

template <int B, int W, int PD>
static int32_t HistAddFullLaunch(hwpa_lane_hist* hist, EdgeLane* edge, const int16_t* traces, uint32_t frames,
                                 const hwpa_lane_model* models, uint32_t lanes, hipStream_t s, bool wrap)
{
    if (frames < (uint32_t)(W * B)) return HistAddLaunch<B, 192>(hist, edge, traces, frames, models, lanes, s, wrap);
    hipLaunchKernelGGL((HistAddFullKernel<B, W, PD>), dim3(lanes), dim3(W * LS), 0, s, hist, edge, traces, (int)frames,
                       models);
    const int32_t rc = CheckLaunch("HistAddFullKernel");
    if (rc != HWPA_OK || !wrap) return rc;
    return HistWrapLaunch(hist, edge, traces, frames, models, lanes, s);
}

// Histogram variants: frames per prefetched batch x bins kept in LDS (x waves per lane).
int32_t HistAddVariant(int32_t variant, hwpa_lane_hist* hist, void* edge, const int16_t* traces, uint32_t frames,
                       const hwpa_lane_model* models, uint32_t lanes, hipStream_t s, bool wrap)
{
    if (!hist || !edge || !traces || !models || lanes == 0) return Fail(HWPA_ERR_INVALID_ARG, "hist_add: bad argument");
    if (reinterpret_cast<uintptr_t>(hist) % 16 != 0)
        return Fail(HWPA_ERR_INVALID_ARG, "hist_add: histograms must be 16-byte aligned");
    EdgeLane* e = (EdgeLane*)edge;
    switch (variant) {
    case 0: return HistAddLaunch<16, 160>(hist, e, traces, frames, models, lanes, s, wrap);
    case 1: return HistAddLaunch<16, 128>(hist, e, traces, frames, models, lanes, s, wrap);
    case 2: return HistAddLaunch<16, 192>(hist, e, traces, frames, models, lanes, s, wrap);
    case 3: return HistAddLaunch<8, 160>(hist, e, traces, frames, models, lanes, s, wrap);
    case 4: return HistAddLaunch<8, 128>(hist, e, traces, frames, models, lanes, s, wrap);
    case 5: return HistAddLaunch<8, 192>(hist, e, traces, frames, models, lanes, s, wrap);
    case 6: return HistAddLaunch<32, 160>(hist, e, traces, frames, models, lanes, s, wrap);
    case 7: return HistAddLaunch<32, 128>(hist, e, traces, frames, models, lanes, s, wrap);
    case 8: return HistAddLaunch<32, 192>(hist, e, traces, frames, models, lanes, s, wrap);
    case 9: return HistAddWavesLaunch<8, 160, 4>(hist, e, traces, frames, models, lanes, s, wrap);
    case 10: return HistAddWavesLaunch<8, 128, 4>(hist, e, traces, frames, models, lanes, s, wrap);
    case 11: return HistAddWavesLaunch<16, 160, 4>(hist, e, traces, frames, models, lanes, s, wrap);
    case 12: return HistAddWavesLaunch<8, 160, 2>(hist, e, traces, frames, models, lanes, s, wrap);
    case 13: return HistAddWavesLaunch<8, 192, 4>(hist, e, traces, frames, models, lanes, s, wrap);
    case 14: return HistAddWavesLaunch<8, 192, 8>(hist, e, traces, frames, models, lanes, s, wrap);
    case 15: return HistAddWavesLaunch<8, 160, 8>(hist, e, traces, frames, models, lanes, s, wrap);
    case 16: return HistAddFullLaunch<8, 4, 4>(hist, e, traces, frames, models, lanes, s, wrap);
    case 17: return HistAddFullLaunch<16, 4, 3>(hist, e, traces, frames, models, lanes, s, wrap);
    case 18: return HistAddFullLaunch<8, 8, 1>(hist, e, traces, frames, models, lanes, s, wrap);
    case 19: return HistAddFullLaunch<8, 8, 2>(hist, e, traces, frames, models, lanes, s, wrap);
    case 20: return HistAddFullLaunch<16, 8, 1>(hist, e, traces, frames, models, lanes, s, wrap);
    case 21: return HistAddFullLaunch<4, 8, 4>(hist, e, traces, frames, models, lanes, s, wrap);
    default: return Fail(HWPA_ERR_INVALID_ARG, "hist_add: unknown variant");
    }
}

int32_t HistAdd(hwpa_lane_hist* hist, void* edge, const int16_t* traces, uint32_t frames,
                const hwpa_lane_model* models, uint32_t lanes, hipStream_t s, bool wrap)
{
    return HistAddVariant(TuneGet(kTuneHistogram, lanes, frames), hist, edge, traces, frames, models, lanes, s, wrap);
}

int32_t HistAddWrap(hwpa_lane_hist* hist, void* edge, const int16_t* traces, uint32_t frames,
                    const hwpa_lane_model* models, uint32_t lanes, hipStream_t s)
{
    if (!hist || !edge || !traces || !models || lanes == 0) return Fail(HWPA_ERR_INVALID_ARG, "hist_add: bad argument");
    return HistWrapLaunch(hist, (EdgeLane*)edge, traces, frames, models, lanes, s);
}

}  // namespace hwpa
