// stats.hip -- baseline-statistics aggregation across chunks
// (BaselineStatsAggregatorHost, traceAnalysis/BaselineStatsAggregatorHost.cpp:40-71;
// merge rules of dataTypes/BaselinerStatAccumulator.cpp:64-75,
// common/AutocorrAccumulator.cpp:72-104, common/StatAccumulator.h:183-191).
// One wave per lane, one thread per ZMW; the autocorrelation buffer indices
// are lane-uniform.

#include "internal.h"

namespace hwpa {

__global__ __launch_bounds__(64) void StatsResetKernel(hwpa_baseliner_stats* agg)
{
    hwpa_baseliner_stats& a = agg[blockIdx.x];
    const int z = threadIdx.x;
    a.full.basic.offset[z] = 0.f;
    a.full.basic.m0[z] = 0.f;
    a.full.basic.m1[z] = 0.f;
    a.full.basic.m2[z] = 0.f;
    a.full.lag_m2[z] = 0.f;
    for (int k = 0; k < 4; ++k) {
        a.full.fbuf[k][z] = 0.f;
        a.full.bbuf[k][z] = 0.f;
    }
    a.full.bidx[0][z] = 0;
    a.full.bidx[1][z] = 0;
    a.trace_min[z] = 32767;
    a.trace_max[z] = -32768;
    a.baseline.offset[z] = 0.f;
    a.baseline.m0[z] = 0.f;
    a.baseline.m1[z] = 0.f;
    a.baseline.m2[z] = 0.f;
    a.raw_baseline_sum[z] = 0.f;
    a.background.offset[z] = 0.f;
    a.background.m0[z] = 0.f;
    a.background.m1[z] = 0.f;
    a.background.m2[z] = 0.f;
}

__global__ __launch_bounds__(64) void StatsMergeKernel(hwpa_baseliner_stats* agg, const hwpa_baseliner_stats* chunk)
{
    hwpa_baseliner_stats& a = agg[blockIdx.x];
    const hwpa_baseliner_stats& o = chunk[blockIdx.x];
    const int z = threadIdx.x;
    // Buffer indices are lane-uniform; clamp so malformed input cannot index out of range.
    const int fbi = min((int)a.full.bidx[0][0], 4), bbi = a.full.bidx[1][0] & 3;
    const int ofbi = min((int)o.full.bidx[0][0], 4), obbi = o.full.bidx[1][0] & 3;
    const int n1 = 4 - ofbi;
    const int n2 = 4 - fbi;
    a.full.basic.m0[z] += o.full.basic.m0[z];
    a.full.basic.m1[z] += o.full.basic.m1[z];
    a.full.basic.m2[z] += o.full.basic.m2[z];
    float lag = a.full.lag_m2[z] + o.full.lag_m2[z];
    float bb[4], fb[4];
    for (int k = 0; k < 4; ++k) {
        bb[k] = a.full.bbuf[k][z];
        fb[k] = a.full.fbuf[k][z];
    }
    for (int k = 0; k < 4 - n1; ++k) {
        lag += bb[(bbi + k) % 4] * o.full.fbuf[k][z];
        bb[(bbi + k) % 4] = o.full.bbuf[(obbi + n1 + k) % 4][z];
    }
    for (int k = 0; k < n2; ++k) fb[fbi + k] = o.full.fbuf[k][z];
    a.full.lag_m2[z] = lag;
    for (int k = 0; k < 4; ++k) {
        a.full.bbuf[k][z] = bb[k];
        a.full.fbuf[k][z] = fb[k];
    }
    const int16_t tmn = a.trace_min[z], otmn = o.trace_min[z];
    const int16_t tmx = a.trace_max[z], otmx = o.trace_max[z];
    a.trace_min[z] = tmn < otmn ? tmn : otmn;
    a.trace_max[z] = tmx > otmx ? tmx : otmx;
    a.baseline.m0[z] += o.baseline.m0[z];
    a.baseline.m1[z] += o.baseline.m1[z];
    a.baseline.m2[z] += o.baseline.m2[z];
    a.raw_baseline_sum[z] += o.raw_baseline_sum[z];
    a.background.m0[z] += o.background.m0[z];
    a.background.m1[z] += o.background.m1[z];
    a.background.m2[z] += o.background.m2[z];
    // All threads read bidx[.][0] above; wait before rewriting it.
    __syncthreads();
    a.full.bidx[0][z] = (uint8_t)(fbi + n2);
    a.full.bidx[1][z] = (uint8_t)((bbi + (4 - n1)) % 4);
}

int32_t StatsReset(hwpa_baseliner_stats* agg, uint32_t lanes, hipStream_t s)
{
    if (!agg || lanes == 0) return Fail(HWPA_ERR_INVALID_ARG, "stats_reset: bad argument");
    hipLaunchKernelGGL(StatsResetKernel, dim3(lanes), dim3(LS), 0, s, agg);
    return CheckLaunch("StatsResetKernel");
}

int32_t StatsMerge(hwpa_baseliner_stats* agg, const hwpa_baseliner_stats* chunk, uint32_t lanes, hipStream_t s)
{
    if (!agg || !chunk || lanes == 0) return Fail(HWPA_ERR_INVALID_ARG, "stats_merge: bad argument");
    hipLaunchKernelGGL(StatsMergeKernel, dim3(lanes), dim3(LS), 0, s, agg, chunk);
    return CheckLaunch("StatsMergeKernel");
}

}  // namespace hwpa
