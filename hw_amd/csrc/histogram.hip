// histogram.hip -- per-ZMW 300-bin trace histograms with edge-frame
// scrubbing (TraceHistogramAccumHost, traceAnalysis/TraceHistogramAccumHost.cpp:92-169;
// UHistogramSimd binning, dataTypes/UHistogramSimd.{h,cpp}; EdgeFrameClassifier,
// traceAnalysis/EdgeFrameClassifier.h).
//
// Mapping: one wave64 per lane, one thread per ZMW.  The lane's 300x64 u16
// counters (38.4 KB) are staged in LDS for the chunk; each thread owns one
// ZMW column, so increments need no atomics, and the [bin][zmw] layout puts
// the 64 lanes of every increment on distinct LDS banks whatever the bins.

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

template <int UNROLL>
__global__ __launch_bounds__(64) void HistAddKernel(hwpa_lane_hist* hist, EdgeLane* edge,
                                                    const int16_t* __restrict__ traces, int frames,
                                                    const hwpa_lane_model* __restrict__ models)
{
    __shared__ uint16_t cnt[NB * LS];
    const int lane = blockIdx.x;
    const int z = threadIdx.x;
    hwpa_lane_hist& h = hist[lane];
    EdgeLane& e = edge[lane];
    for (int b = 0; b < NB; ++b) cnt[b * LS + z] = h.bin_count[b][z];
    uint32_t lowOut = h.outlier_low[z], highOut = h.outlier_high[z];
    const float lo = h.low_bound[z];
    const float bs = h.bin_size[z];

    // Edge threshold from the current model (TraceHistogramAccumHost.cpp:156-157)
    const hwpa_lane_model& m = models[lane];
    const float bvar = hw_h2f(m.baseline.vars[z]);
    const float bmean = hw_h2f(m.baseline.means[z]);
    const int thr = hw_wrap16(hw_round_i32(2.0f * hw_sqrtf(bvar) + bmean));

    int prev = e.prev[z];
    bool h0 = e.hist[0][z] != 0, h1 = e.hist[1][z] != 0;
    int i0 = e.i0, i1 = e.i1;
    __syncthreads();

    const int16_t* tr = traces + (size_t)lane * frames * LS + z;
    for (int t0 = 0; t0 < frames; t0 += UNROLL) {
        int fv[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) fv[u] = (t0 + u < frames) ? (int)tr[(size_t)(t0 + u) * LS] : 0;
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) {
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
                const float y = ((float)cand - lo) / bs;
                const int bin = hw_floor_i32(y);
                if (bin < 0) lowOut = (lowOut + 1) & 0xffffu;
                else if (bin >= NB) highOut = (highOut + 1) & 0xffffu;
                else cnt[bin * LS + z] = (uint16_t)(cnt[bin * LS + z] + 1);
            }
        }
    }
    __syncthreads();
    for (int b = 0; b < NB; ++b) h.bin_count[b][z] = cnt[b * LS + z];
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
    hipLaunchKernelGGL(HistAddKernel<8>, dim3(lanes), dim3(LS), 0, s, hist, (EdgeLane*)edge, traces, (int)frames,
                       models);
    return CheckLaunch("HistAddKernel");
}

}  // namespace hwpa
