// histogram.hip -- per-ZMW 300-bin trace histograms with edge-frame
// scrubbing (TraceHistogramAccumHost, traceAnalysis/TraceHistogramAccumHost.cpp:92-169;
// UHistogramSimd binning, dataTypes/UHistogramSimd.{h,cpp}; EdgeFrameClassifier,
// traceAnalysis/EdgeFrameClassifier.h).
//
// Mapping: one wave64 per lane, one thread per ZMW.  The lane's 300x64 u16
// counters (38.4 KB) are staged in LDS for the chunk; each thread owns one
// ZMW column, and the [bin pair][zmw] layout puts the 64 lanes of every
// increment on distinct LDS banks whatever the bins.  Frames are read in
// batches of 16 with the next batch in flight while one is binned.

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
    int32_t pad[30];
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
__global__ __launch_bounds__(64) void HistResetKernel(hwpa_lane_hist* hist, const hwpa_baseliner_stats* stats,
                                                      HistConsts k)
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
    for (int b = 0; b < NB; ++b) h.bin_count[b][z] = 0;
}

// Bin of a sample: floor((x - lo) / bs) with the reference's IEEE divide
// (UHistogramSimd::BinIndex, dataTypes/UHistogramSimd.h:199-229).  The
// quotient is first formed as (x - lo) * (1 / bs), which is within
// 3 * 2^-24 |y| of the divided value; for |y| < 2048 that is below 3.7e-4, so
// when the product is more than 1/1024 away from an integer both share the
// same floor, and only near a bin edge is the divide itself evaluated.  For
// 2048 <= |y| < 2^30 the sample is an outlier on the same side either way;
// NaN, infinities and huge quotients take the divide (whose float -> int
// conversion gives 0x80000000 out of range, a low outlier, as on the host).
__device__ __forceinline__ int BinOf(float a, float bs, float inv)
{
    const float yp = a * inv;
    const float ay = hw_fabsf(yp);
    if (ay >= 2048.0f && ay < 1073741824.0f) return (yp > 0.0f) ? NB : -1;
    if (ay < 2048.0f) {
        const float fl = hw_floorf(yp);
        const float d = yp - fl;
        if (d >= 0.0009765625f && d <= 0.9990234375f) return (int)fl;
    }
    return hw_floor_i32(a / bs);
}

// TraceHistogramAccumHost::AddBlock (TraceHistogramAccumHost.cpp:128-169).
// Counters live in LDS as words holding bins 2j (low half) and 2j+1 (high
// half) of one ZMW, so an increment is a single ds_add_u32 that nothing waits
// on.  That cannot carry into the neighbouring half while every count stays
// below 65536 -- checked before the chunk; otherwise the wave increments with
// a 16-bit wrapping read-modify-write, the reference's u16 counter semantics.
template <int B>
__global__ __launch_bounds__(64) void HistAddKernel(hwpa_lane_hist* hist, EdgeLane* edge,
                                                    const int16_t* __restrict__ traces, int frames,
                                                    const hwpa_lane_model* __restrict__ models)
{
    __shared__ uint32_t cnt[(NB / 2) * LS];
    const int lane = blockIdx.x;
    const int z = threadIdx.x;
    hwpa_lane_hist& h = hist[lane];
    EdgeLane& e = edge[lane];
    uint32_t mx = 0;
    for (int j = 0; j < NB / 2; ++j) {
        const uint32_t c0 = h.bin_count[2 * j][z], c1 = h.bin_count[2 * j + 1][z];
        mx = c0 > mx ? c0 : mx;
        mx = c1 > mx ? c1 : mx;
        cnt[j * LS + z] = c0 | (c1 << 16);
    }
    const bool fast = __all(mx + (uint32_t)frames <= 65535u);
    uint32_t lowOut = h.outlier_low[z], highOut = h.outlier_high[z];
    const float lo = h.low_bound[z];
    const float bs = h.bin_size[z];
    const float inv = 1.0f / bs;

    // Edge threshold from the current model (TraceHistogramAccumHost.cpp:156-157)
    const hwpa_lane_model& m = models[lane];
    const float bvar = hw_h2f(m.baseline.vars[z]);
    const float bmean = hw_h2f(m.baseline.means[z]);
    const int thr = hw_wrap16(hw_round_i32(2.0f * hw_sqrtf(bvar) + bmean));

    int prev = e.prev[z];
    bool h0 = e.hist[0][z] != 0, h1 = e.hist[1][z] != 0;
    int i0 = e.i0, i1 = e.i1;

    const int16_t* tr = traces + (size_t)lane * frames * LS + z;
    int fv[B];
#pragma unroll
    for (int u = 0; u < B; ++u) fv[u] = (u < frames) ? (int)tr[(size_t)u * LS] : 0;
    for (int t0 = 0; t0 < frames; t0 += B) {
        int nv[B];
        const int tn = t0 + B;
#pragma unroll
        for (int u = 0; u < B; ++u) nv[u] = (tn + u < frames) ? (int)tr[(size_t)(tn + u) * LS] : 0;
#pragma unroll
        for (int u = 0; u < B; ++u) {
            if (t0 + u >= frames) break;
            const int f = fv[u];
            const bool isBl = f < thr;
            // CircularArray<LaneMask,2>: full when two entries have been pushed
            const bool full = (i1 >= 0) && (((2 + i1 - i0) & 1) + 1) == 2;
            const bool front = (i0 == 0) ? h0 : h1;
            const bool isEdge = full ? (isBl != front) : true;
            if (full) i0 = (i0 + 1) & 1;
            i1 = (i1 + 1) & 1;
            if (i1 == 0) h0 = isBl;
            else h1 = isBl;
            const int cand = prev;
            prev = f;
            if (!isEdge) {
                const int bin = BinOf((float)cand - lo, bs, inv);
                if (bin < 0) {
                    lowOut = (lowOut + 1) & 0xffffu;
                } else if (bin >= NB) {
                    highOut = (highOut + 1) & 0xffffu;
                } else {
                    uint32_t* w = &cnt[(bin >> 1) * LS + z];
                    const uint32_t sh = (uint32_t)(bin & 1) * 16u;
                    if (fast) {
                        atomicAdd(w, 1u << sh);
                    } else {
                        const uint32_t v = *w;
                        const uint32_t c = ((v >> sh) + 1u) & 0xffffu;
                        *w = (v & ~(0xffffu << sh)) | (c << sh);
                    }
                }
            }
        }
#pragma unroll
        for (int u = 0; u < B; ++u) fv[u] = nv[u];
    }
    for (int j = 0; j < NB / 2; ++j) {
        const uint32_t v = cnt[j * LS + z];
        h.bin_count[2 * j][z] = (uint16_t)(v & 0xffffu);
        h.bin_count[2 * j + 1][z] = (uint16_t)(v >> 16);
    }
    h.outlier_low[z] = (uint16_t)lowOut;
    h.outlier_high[z] = (uint16_t)highOut;
    e.prev[z] = (int16_t)prev;
    e.hist[0][z] = h0 ? 1 : 0;
    e.hist[1][z] = h1 ? 1 : 0;
    if (z == 0) {
        e.i0 = i0;
        e.i1 = i1;
    }
}

int32_t HistReset(const hwpa_config& cfg, hwpa_lane_hist* hist, const hwpa_baseliner_stats* stats, uint32_t lanes,
                  hipStream_t s)
{
    if (!hist || !stats || lanes == 0) return Fail(HWPA_ERR_INVALID_ARG, "hist_reset: bad argument");
    HistConsts k;
    k.binSizeCoeff = cfg.bin_size_coeff;
    k.binSizeMin = cfg.bin_size_low_bound_coeff * std::max(1.0f, cfg.movie.photoelectron_sensitivity);
    k.fallbackSigma = cfg.fallback_baseline_sigma;
    k.minFrameCount = (float)cfg.baseline_stat_min_frame_count;
    hipLaunchKernelGGL(HistResetKernel, dim3(lanes), dim3(LS), 0, s, hist, stats, k);
    return CheckLaunch("HistResetKernel");
}

int32_t HistAdd(hwpa_lane_hist* hist, void* edge, const int16_t* traces, uint32_t frames,
                const hwpa_lane_model* models, uint32_t lanes, hipStream_t s)
{
    if (!hist || !edge || !traces || !models || lanes == 0) return Fail(HWPA_ERR_INVALID_ARG, "hist_add: bad argument");
    hipLaunchKernelGGL(HistAddKernel<16>, dim3(lanes), dim3(LS), 0, s, hist, (EdgeLane*)edge, traces, (int)frames,
                       models);
    return CheckLaunch("HistAddKernel");
}

}  // namespace hwpa
