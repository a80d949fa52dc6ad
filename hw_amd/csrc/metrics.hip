// metrics.hip -- HF metrics: per-ZMW basecalling metrics accumulated over
// blocks of framesPerHFMetricBlock frames (HFMetricsFilterHost,
// traceAnalysis/HFMetricsFilterHost.cpp:37-116; BasecallingMetricsAccumulator,
// dataTypes/BasecallingMetricsAccumulator.cpp:62-504; TraceAnalysisMetrics,
// dataTypes/TraceAnalysisMetrics.h:47-190).
//
// Mapping: one wave64 per lane, one thread per ZMW; the accumulator of a
// lane is a structure of 64-wide rows in HBM.  Each chunk one kernel counts
// the ZMW's pulses (its own slot list), keeps the latest model and merges the
// chunk's baseliner autocorrelation and pulse-detector baseline statistics;
// the lane-wide decisions of the reference's LaneArray code (all()/any()) are
// wave votes.  At the end of a block a second kernel finalises the moments,
// evaluates the activity-labeler tree and writes BasecallingMetrics.

#include "internal.h"

#include <cmath>
#include <cstring>
#include <new>
#include <vector>

namespace hwpa {

constexpr uint8_t kLabelNone = 5;  // Pulse::NucleotideLabel::NONE (dataTypes/Pulse.h:52-57)
constexpr int kTreeFeatures = 11;  // ActivityLabeler::NUM_FEATURES (dataTypes/TrainedCartParams.h:57-70)
constexpr int32_t kMaxTreeNodes = 65536;

struct HfmLaneDev {
    uint16_t numPulseFrames[LS], numBaseFrames[LS], numSandwiches[LS], numHalfSandwiches[LS], numStutters[LS];
    uint8_t activity[LS];
    uint8_t cLabel[2][LS];  // basecall caches: [0] previous pulse, [1] the one before
    uint16_t cWidth[2][LS];
    uint32_t cStart[2][LS];
    float pkMidSignal[NA][LS], bpZvar[NA][LS], pkZvar[NA][LS], pkMax[NA][LS], modelVar[NA][LS], modelMean[NA][LS];
    uint16_t numPkMidFrames[NA][LS], numPkMidBases[NA][LS], numBasesBy[NA][LS], numPulsesBy[NA][LS];
    uint32_t startFrame[LS], numFrames[LS];
    float pdScore[LS];
    float blOff[LS], blM0[LS], blM1[LS], blM2[LS];  // baselineStatAccum_ (offset survives Reset)
    float acOff[LS], acM0[LS], acM1[LS], acM2[LS], acLag[LS], acF[4][LS], acB[4][LS];  // autocorrAccum_
    int32_t fbi, bbi;                                                                    // lane-wide
    int32_t pad[14];
};

struct TreeDev {
    int32_t nodes = 0;
    float maxHsw = 0.0f, hsw[2] = {0.0f, 0.0f};
    const int16_t* left = nullptr;
    const int16_t* right = nullptr;
    const int8_t* feature = nullptr;
    const float* threshold = nullptr;
    const int8_t* value = nullptr;
};

__device__ __forceinline__ bool IsNan(float x) { return x != x; }

// StatAccumulator::Mean / Variance (common/StatAccumulator.h:91-105)
__device__ __forceinline__ float StatMean(float m0, float m1, float off) { return m1 / m0 + off; }
__device__ __forceinline__ float StatVariance(float m0, float m1, float m2)
{
    float var = m1 * m1 / m0;
    var = (m2 - var) / (m0 - 1.0f);
    var = hw_maxf(var, 0.0f);
    return (m0 > 1.0f) ? var : hw_u2f(0x7fc00000u);
}

// AutocorrAccumulator::Autocorrelation (common/AutocorrAccumulator.cpp:42-67)
__device__ float Autocorrelation(float off, float m0, float m1, float m2, float lag, const float* f,
                                 const float* b)
{
    const float nmk = m0 - 4.0f;
    const float mu = StatMean(m0, m1, off) - off;
    float m1x2 = 2.0f * m1;
    for (int k = 0; k < 4; ++k) m1x2 -= f[k] + b[k];
    float ac = mu * (m1x2 - nmk * mu);
    ac = (lag - ac) / ((nmk - 1.0f) * StatVariance(m0, m1, m2));
    ac = hw_minf(hw_maxf(ac, -1.0f), 1.0f);
    return (nmk < 1.0f || IsNan(ac)) ? hw_u2f(0x7fc00000u) : ac;
}

// StatAccumulator::Offset(value) (common/StatAccumulator.h:234-246); the early
// return is lane-wide.
__device__ __forceinline__ void SetOffset(float& off, float m0, float& m1, float& m2, float value)
{
    if (__all(value == off)) return;
    const float m1new = m1 + m0 * (off - value);
    const float m2new = m2 + (m1new * m1new - m1 * m1) / m0;
    off = value;
    m1 = m1new;
    m2 = (m0 == 0.0f) ? 0.0f : m2new;
}

// One chunk: [Reset,] Count, AddModels, AddBatchMetrics (BasecallingMetricsAccumulator.cpp:254-331).
__global__ __launch_bounds__(64) void HfmChunkKernel(HfmLaneDev* state, const hwpa_pulse* __restrict__ pulses,
                                                     const uint32_t* __restrict__ counts, uint32_t maxCalls,
                                                     uint32_t frames, const hwpa_baseliner_stats* __restrict__ blStats,
                                                     const float* __restrict__ score,
                                                     const hwpa_stat_acc* __restrict__ pdStats,
                                                     const hwpa_lane_model* __restrict__ models, int32_t reset)
{
    const int lane = blockIdx.x;
    const int z = threadIdx.x;
    HfmLaneDev& S = state[lane];
    int fbi = S.fbi, bbi = S.bbi;

    if (reset) {
        S.numPulseFrames[z] = S.numBaseFrames[z] = S.numSandwiches[z] = S.numHalfSandwiches[z] = S.numStutters[z] = 0;
        S.activity[z] = 0;
        for (int c = 0; c < 2; ++c) {
            S.cStart[c][z] = 0;
            S.cWidth[c][z] = 0;
            S.cLabel[c][z] = kLabelNone;
        }
        for (int a = 0; a < NA; ++a) {
            S.pkMidSignal[a][z] = S.bpZvar[a][z] = S.pkZvar[a][z] = S.pkMax[a][z] = 0.0f;
            S.numPkMidFrames[a][z] = S.numPkMidBases[a][z] = S.numBasesBy[a][z] = S.numPulsesBy[a][z] = 0;
        }
        S.blM0[z] = S.blM1[z] = S.blM2[z] = 0.0f;
        S.startFrame[z] = S.startFrame[z] + S.numFrames[z];
        S.numFrames[z] = 0;
        S.pdScore[z] = 0.0f;
        S.acM0[z] = S.acM1[z] = S.acM2[z] = S.acLag[z] = 0.0f;
        for (int k = 0; k < 4; ++k) S.acF[k][z] = S.acB[k][z] = 0.0f;
        fbi = bbi = 0;
    }

    // Count (BasecallingMetricsAccumulator.cpp:333-410); labels beyond T are not counted per analog.
    // The per-analog accumulators live in registers for the chunk (selected by compare, so the
    // label never indexes memory) and each keeps the reference's order of additions.
    {
        S.numFrames[z] += frames;
        uint32_t pS = S.cStart[0][z], ppS = S.cStart[1][z];
        uint32_t pW = S.cWidth[0][z], ppW = S.cWidth[1][z];
        uint32_t pL = S.cLabel[0][z], ppL = S.cLabel[1][z];
        uint32_t nPF = S.numPulseFrames[z], nBF = S.numBaseFrames[z], nSw = S.numSandwiches[z];
        uint32_t nHs = S.numHalfSandwiches[z], nSt = S.numStutters[z];
        uint32_t cP[NA], cB[NA], cMB[NA], cMF[NA];
        float aMid[NA], aBp[NA], aPk[NA], aMax[NA];
#pragma unroll
        for (int a = 0; a < NA; ++a) {
            cP[a] = S.numPulsesBy[a][z];
            cB[a] = S.numBasesBy[a][z];
            cMB[a] = S.numPkMidBases[a][z];
            cMF[a] = S.numPkMidFrames[a][z];
            aMid[a] = S.pkMidSignal[a][z];
            aBp[a] = S.bpZvar[a][z];
            aPk[a] = S.pkZvar[a][z];
            aMax[a] = S.pkMax[a][z];
        }
        const size_t zg = (size_t)lane * LS + z;
        const uint32_t n = counts[zg] < maxCalls ? counts[zg] : maxCalls;
        // The 20-byte record as five dwords (records are 4-byte aligned).
        const uint32_t* zp = reinterpret_cast<const uint32_t*>(pulses + zg * maxCalls);
        uint32_t r[5] = {0u, 0u, 0u, 0u, 0u};
        if (n > 0)
#pragma unroll
            for (int k = 0; k < 5; ++k) r[k] = zp[k];
        for (uint32_t bi = 0; bi < n; ++bi) {
            uint32_t nx[5] = {0u, 0u, 0u, 0u, 0u};
            if (bi + 1 < n)
#pragma unroll
                for (int k = 0; k < 5; ++k) nx[k] = zp[5 * (bi + 1) + k];
            const uint32_t start = r[0];
            const uint32_t w = r[1] & 0xffffu;
            const uint32_t midS = r[2] & 0xffffu;
            const uint32_t maxS = r[2] >> 16;
            const float m2 = __uint_as_float(r[3]);
            const uint32_t label = r[4] & 0xffu;
            const bool reject = ((r[4] >> 8) & 0xffu) != 0u;
            nPF = (nPF + w) & 0xffffu;
            const float mx = (float)maxS * 0.1f;
            if (pL != kLabelNone) {
                if (pL == label) nSt = (nSt + 1) & 0xffffu;
                const bool abutted = start == pS + pW;
                if (abutted && pL != label) nHs = (nHs + 1) & 0xffffu;
                if (ppL != kLabelNone) {
                    const bool prevAbutted = pS == ppS + ppW;
                    if (prevAbutted && abutted && ppL == label && ppL != pL) nSw = (nSw + 1) & 0xffffu;
                }
            }
            if (!reject) nBF = (nBF + w) & 0xffffu;
            const bool mid3 = !reject && w >= 3u;  // MidSignal() is NaN below 3 frames
            const uint32_t midWidth = (w - 2u) & 0xffffu;
            const float mid = (float)midS * 0.1f;
            const float fMidW = (float)midWidth;
#pragma unroll
            for (int a = 0; a < NA; ++a) {
                if (label == (uint32_t)a) {
                    cP[a] = (cP[a] + 1u) & 0xffffu;
                    aMax[a] = (aMax[a] < mx) ? mx : aMax[a];
                    if (!reject) cB[a] = (cB[a] + 1u) & 0xffffu;
                    if (mid3) {
                        cMB[a] = (cMB[a] + 1u) & 0xffffu;
                        cMF[a] = (cMF[a] + midWidth) & 0xffffu;
                        aMid[a] = aMid[a] + mid * fMidW;
                        aBp[a] = aBp[a] + mid * mid * fMidW;
                        aPk[a] = aPk[a] + m2;
                    }
                }
            }
            ppS = pS;
            ppW = pW;
            ppL = pL;
            pS = start;
            pW = w;
            pL = label;
#pragma unroll
            for (int k = 0; k < 5; ++k) r[k] = nx[k];
        }
#pragma unroll
        for (int a = 0; a < NA; ++a) {
            S.numPulsesBy[a][z] = (uint16_t)cP[a];
            S.numBasesBy[a][z] = (uint16_t)cB[a];
            S.numPkMidBases[a][z] = (uint16_t)cMB[a];
            S.numPkMidFrames[a][z] = (uint16_t)cMF[a];
            S.pkMidSignal[a][z] = aMid[a];
            S.bpZvar[a][z] = aBp[a];
            S.pkZvar[a][z] = aPk[a];
            S.pkMax[a][z] = aMax[a];
        }
        S.numPulseFrames[z] = (uint16_t)nPF;
        S.numBaseFrames[z] = (uint16_t)nBF;
        S.numSandwiches[z] = (uint16_t)nSw;
        S.numHalfSandwiches[z] = (uint16_t)nHs;
        S.numStutters[z] = (uint16_t)nSt;
        S.cStart[0][z] = pS;
        S.cWidth[0][z] = (uint16_t)pW;
        S.cLabel[0][z] = (uint8_t)pL;
        S.cStart[1][z] = ppS;
        S.cWidth[1][z] = (uint16_t)ppW;
        S.cLabel[1][z] = (uint8_t)ppL;
    }

    // AddModels (BasecallingMetricsAccumulator.cpp:302-313): the latest model only.
    const hwpa_lane_model& m = models[lane];
#pragma unroll
    for (int a = 0; a < NA; ++a) {
        S.modelVar[a][z] = hw_h2f(m.analogs[a].vars[z]);
        S.modelMean[a][z] = hw_h2f(m.analogs[a].means[z]);
    }

    // AddBatchMetrics (BasecallingMetricsAccumulator.cpp:315-331)
    const hwpa_baseliner_stats& bs = blStats[lane];
    {
        // AutocorrAccumulator::Merge (common/AutocorrAccumulator.cpp:72-102), this on the left.
        const hwpa_autocorr_acc& o = bs.full;
        // A valid state has 0 <= fbi <= 4 and 0 <= bbi < 4; the clamps keep every index in bounds.
        const int ofbi = o.bidx[0][0] < 4 ? (int)o.bidx[0][0] : 4, obbi = o.bidx[1][0] & 3;
        const int n1 = 4 - ofbi, n2 = 4 - fbi;
        float f[4], b[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            f[k] = S.acF[k][z];
            b[k] = S.acB[k][z];
        }
        float lag = S.acLag[z] + o.lag_m2[z];
        S.acM0[z] = S.acM0[z] + o.basic.m0[z];
        S.acM1[z] = S.acM1[z] + o.basic.m1[z];
        S.acM2[z] = S.acM2[z] + o.basic.m2[z];
        for (int k = 0; k < 4 - n1; ++k) {
            const int bi = (bbi + k) % 4;
            float bv = 0.0f;
#pragma unroll
            for (int q = 0; q < 4; ++q) bv = (q == bi) ? b[q] : bv;
            float of = 0.0f;
#pragma unroll
            for (int q = 0; q < 4; ++q) of = (q == k) ? o.fbuf[q][z] : of;
            lag += bv * of;
            const float nb = o.bbuf[(obbi + n1 + k) % 4][z];
#pragma unroll
            for (int q = 0; q < 4; ++q) b[q] = (q == bi) ? nb : b[q];
        }
        for (int k = 0; k < n2; ++k) {
            const float nf = o.fbuf[k][z];
#pragma unroll
            for (int q = 0; q < 4; ++q) f[q] = (q == fbi + k) ? nf : f[q];
        }
        S.acLag[z] = lag;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            S.acF[k][z] = f[k];
            S.acB[k][z] = b[k];
        }
        fbi += n2;
        bbi = (bbi + (4 - n1)) % 4;
    }
    S.pdScore[z] = S.pdScore[z] + score[(size_t)lane * LS + z];
    {
        // Pulse-detector baseline stats shifted by the background mean, then
        // StatAccumulator::operator+= (common/StatAccumulator.h:145-170).
        const hwpa_stat_acc& pd = pdStats[lane];
        const float bgMean = StatMean(bs.background.m0[z], bs.background.m1[z], bs.background.offset[z]);
        float pOff = pd.offset[z] + bgMean, pM0 = pd.m0[z], pM1 = pd.m1[z], pM2 = pd.m2[z];
        float off = S.blOff[z], m0 = S.blM0[z], m1 = S.blM1[z], m2 = S.blM2[z];
        if (!__all(pM0 == 0.0f)) {
            if (__any(IsNan(off))) {
                if (IsNan(off)) {
                    m0 = m1 = m2 = 0.0f;
                    off = 0.0f;
                }
            }
            const float w = pM0 / (m0 + pM0);
            const float on = (1.0f - w) * off + w * pOff;
            const float offNew = IsNan(on) ? off : on;
            SetOffset(pOff, pM0, pM1, pM2, offNew);
            SetOffset(off, m0, m1, m2, offNew);
            m0 += pM0;
            m1 += pM1;
            m2 += pM2;
        }
        S.blOff[z] = off;
        S.blM0[z] = m0;
        S.blM1[z] = m1;
        S.blM2[z] = m2;
    }
    if (z == 0) {
        S.fbi = fbi;
        S.bbi = bbi;
    }
}

// LaneArray sqrt (common/simd/m512f_SSE.h:406-415): 0 maps to 0.
__device__ __forceinline__ float LaneSqrt(float x) { return (x == 0.0f) ? 0.0f : hw_sqrtf(x); }
__device__ __forceinline__ float NanToZero(float x) { return IsNan(x) ? 0.0f : x; }

// FinalizeMetrics (:412-500), LabelBlock (:62-200) and PopulateBasecallingMetrics (:202-252).
// `end` = 0 makes the launch a no-op: the kernel is enqueued every chunk so that a
// captured chunk step has the same launch sequence whether or not a block ends.
__global__ __launch_bounds__(64) void HfmFinalizeKernel(HfmLaneDev* state, TreeDev tree, int32_t realtime,
                                                        float frameRate, hwpa_basecalling_metrics* out, int32_t end)
{
    if (!end) return;
    const int lane = blockIdx.x;
    const int z = threadIdx.x;
    HfmLaneDev& S = state[lane];
    const float nan = hw_u2f(0x7fc00000u);
    const float blVar = StatVariance(S.blM0[z], S.blM1[z], S.blM2[z]);
    float pkMid[NA], bpZ[NA], pkZ[NA];
#pragma unroll
    for (int a = 0; a < NA; ++a) {
        const uint16_t nbU = S.numPkMidBases[a][z], nfU = S.numPkMidFrames[a][z];
        float pkM = S.pkMidSignal[a][z], bp = S.bpZvar[a][z], pk = S.pkZvar[a][z];
        if (__all((int16_t)nbU == 0)) {
            pkM = bp = pk = nan;
        } else if (__all((float)nbU < 2.0f) || __all((float)nfU < 2.0f)) {
            bp = pk = nan;
        } else {
            const float nf = (float)nfU, nb = (float)nbU;
            bp = (bp - (pkM * pkM / nf)) / nf;
            bp = bp * nb / (nb - 1.0f);
            bp -= blVar / (nf / nb);
            pk = (pk - (pkM * pkM / nf)) / (nf - 1.0f);
            pk -= bp + blVar;
            const float miv = S.modelVar[a][z];
            bool mask = miv < blVar;
            pk /= miv - blVar;
            mask = mask || (pk < 0.0f);
            pk = mask ? 0.0f : pk;
            const float mid = pkM / nf;
            mask = mid < 0.0f;
            bp /= (mid * mid);
            mask = mask || (bp < 0.0f);
            bp = mask ? nan : bp;
            const int16_t nbi = (int16_t)nbU;
            mask = nbi == 0;
            pkM = mask ? nan : pkM;
            mask = mask || (nbi < 2);
            mask = mask || ((int16_t)nfU < 2);
            bp = mask ? nan : bp;
            pk = mask ? nan : pk;
        }
        pkMid[a] = pkM;
        bpZ[a] = bp;
        pkZ[a] = pk;
    }
    const uint32_t numFrames = S.numFrames[z];
    const float pdScore = NanToZero(S.pdScore[z] / (float)numFrames);

    float fAc[4], bAc[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        fAc[k] = S.acF[k][z];
        bAc[k] = S.acB[k][z];
    }
    const float acr = Autocorrelation(S.acOff[z], S.acM0[z], S.acM1[z], S.acM2[z], S.acLag[z], fAc, bAc);

    uint32_t nPulses = 0, nBases = 0;
#pragma unroll
    for (int a = 0; a < NA; ++a) {
        nPulses = (nPulses + S.numPulsesBy[a][z]) & 0xffffu;
        nBases = (nBases + S.numBasesBy[a][z]) & 0xffffu;
    }

    uint8_t activity = S.activity[z];
    if (realtime && tree.nodes > 0) {
        float feat[kTreeFeatures];
#pragma unroll
        for (int i = 0; i < kTreeFeatures; ++i) feat[i] = 0.0f;
        const float stdDev = LaneSqrt(blVar);
        const float fB = (float)nBases, fP = (float)nPulses;
        const float pulseWidth = NanToZero((float)S.numPulseFrames[z] / fP);
        float pkmean[NA];
#pragma unroll
        for (int a = 0; a < NA; ++a) pkmean[a] = pkMid[a] / (float)S.numPkMidFrames[a][z];
        const float seconds = (float)numFrames / frameRate;
        feat[0] = fP / seconds;
        feat[1] = NanToZero((float)S.numSandwiches[z] / fP);
        const float hswr = NanToZero((float)S.numHalfSandwiches[z] / fP);
        float hswrExp = tree.hsw[0] * feat[0] + tree.hsw[1];
        hswrExp = (hswrExp > tree.maxHsw) ? tree.maxHsw : hswrExp;
        feat[2] = hswr - hswrExp;
        feat[3] = pdScore;
        feat[4] = pulseWidth;
        feat[5] = NanToZero((float)S.numStutters[z] / fP);
        float rel[NA];
#pragma unroll
        for (int a = 0; a < NA; ++a) rel[a] = S.modelMean[a][z];
        float maxamp = rel[0];
#pragma unroll
        for (int a = 1; a < NA; ++a) maxamp = hw_maxf(rel[a], maxamp);
#pragma unroll
        for (int a = 0; a < NA; ++a) rel[a] /= maxamp;
        int lowIdx = 0;
        float minamp = rel[0];
#pragma unroll
        for (int a = 1; a < NA; ++a) {
            lowIdx = (rel[a] < minamp) ? a : lowIdx;
            minamp = hw_minf(rel[a], minamp);
        }
#pragma unroll
        for (int a = 0; a < NA; ++a) feat[6] += pkmean[a] / stdDev * (float)S.numBasesBy[a][z] / fB * minamp / rel[a];
        feat[6] = NanToZero(feat[6]);
#pragma unroll
        for (int a = 0; a < NA; ++a) feat[7] = hw_maxf(feat[7], (S.pkMax[a][z] - pkmean[a]) / stdDev);
        feat[7] = NanToZero(feat[7]);
        feat[8] = acr;
#pragma unroll
        for (int a = 0; a < NA; ++a) {
            feat[9] += (IsNan(bpZ[a]) || lowIdx == a) ? 0.0f : bpZ[a];
            feat[10] += (IsNan(pkZ[a]) || lowIdx == a) ? 0.0f : pkZ[a];
        }
        feat[9] /= 3.0f;
        feat[10] /= 3.0f;
        // Tree walk: children have larger indices than their parent, so at most `nodes` steps.
        int cur = 0;
        for (int step = 0; step < tree.nodes; ++step) {
            const int fi = tree.feature[cur];
            if (fi < 0) break;
            float fv = 0.0f;
#pragma unroll
            for (int i = 0; i < kTreeFeatures; ++i) fv = (i == fi) ? feat[i] : fv;
            cur = (fv <= tree.threshold[cur]) ? tree.left[cur] : tree.right[cur];
        }
        activity = (uint8_t)tree.value[cur];
    }

    hwpa_basecalling_metrics& o = out[lane];
    o.activity_label[z] = activity;
    o.num_pulse_frames[z] = S.numPulseFrames[z];
    o.num_base_frames[z] = S.numBaseFrames[z];
    o.num_sandwiches[z] = S.numSandwiches[z];
    o.num_half_sandwiches[z] = S.numHalfSandwiches[z];
    o.num_pulse_label_stutters[z] = S.numStutters[z];
    o.num_bases[z] = (uint16_t)nBases;
    o.num_pulses[z] = (uint16_t)nPulses;
#pragma unroll
    for (int a = 0; a < NA; ++a) {
        o.pk_mid_signal[a][z] = pkMid[a];
        o.bp_zvar[a][z] = bpZ[a];
        o.pk_zvar[a][z] = pkZ[a];
        o.pk_max[a][z] = S.pkMax[a][z];
        o.num_pk_mid_frames[a][z] = S.numPkMidFrames[a][z];
        o.num_pk_mid_bases_by_analog[a][z] = S.numPkMidBases[a][z];
        o.num_bases_by_analog[a][z] = S.numBasesBy[a][z];
        o.num_pulses_by_analog[a][z] = S.numPulsesBy[a][z];
    }
    o.frame_baseline_dws[z] = StatMean(S.blM0[z], S.blM1[z], S.blOff[z]);
    o.frame_baseline_variance_dws[z] = blVar;
    o.num_frames_baseline[z] = (uint16_t)hw_round_i32(S.blM0[z]);
    o.start_frame[z] = S.startFrame[z];
    o.num_frames[z] = (uint16_t)numFrames;
    o.autocorrelation[z] = NanToZero(acr);
    o.pulse_detection_score[z] = pdScore;
    o.pixel_checksum[z] = 0;
    // The finalised values stay in the state only as far as the reference keeps them: the
    // next chunk starts a new block with a reset.
    S.activity[z] = activity;
}

__global__ __launch_bounds__(64) void HfmInitKernel(HfmLaneDev* state)
{
    HfmLaneDev& S = state[blockIdx.x];
    const int z = threadIdx.x;
    // TraceAnalysisMetrics ctor (TraceAnalysisMetrics.h:56-64) and AddModels-before-use defaults.
    S.startFrame[z] = 0;
    S.numFrames[z] = 0;
    S.pdScore[z] = 0.0f;
    S.blOff[z] = S.blM0[z] = S.blM1[z] = S.blM2[z] = 0.0f;
    S.acOff[z] = S.acM0[z] = S.acM1[z] = S.acM2[z] = S.acLag[z] = 0.0f;
    for (int k = 0; k < 4; ++k) S.acF[k][z] = S.acB[k][z] = 0.0f;
    for (int a = 0; a < NA; ++a) S.modelVar[a][z] = S.modelMean[a][z] = 0.0f;
    if (z == 0) {
        S.fbi = 0;
        S.bbi = 0;
    }
}

}  // namespace hwpa

using namespace hwpa;

struct hwpa_metrics {
    hwpa_config cfg;
    uint32_t lanes = 0;
    uint32_t framesSeen = 0;
    HfmLaneDev* state = nullptr;
    TreeDev tree;
    void* treeMem = nullptr;
};

namespace hwpa {

int32_t MetricsCreate(const hwpa_config& cfg, uint32_t lanes, hwpa_metrics** out)
{
    *out = nullptr;
    if (lanes == 0) return Fail(HWPA_ERR_INVALID_ARG, "metrics: no lanes");
    auto* m = new (std::nothrow) hwpa_metrics();
    if (!m) return Fail(HWPA_ERR_OOM, "out of host memory");
    m->cfg = cfg;
    m->lanes = lanes;
    int32_t rc = DeviceAlloc((void**)&m->state, sizeof(HfmLaneDev) * lanes, "metrics state");
    if (rc != HWPA_OK) {
        delete m;
        return rc;
    }
    hipLaunchKernelGGL(HfmInitKernel, dim3(lanes), dim3(LS), 0, nullptr, m->state);
    if ((rc = CheckLaunch("HfmInitKernel")) != HWPA_OK || hipDeviceSynchronize() != hipSuccess) {
        MetricsDestroy(m);
        return rc != HWPA_OK ? rc : Fail(HWPA_ERR_HIP, "metrics init failed");
    }
    *out = m;
    return HWPA_OK;
}

void MetricsDestroy(hwpa_metrics* m)
{
    if (!m) return;
    DeviceFree(m->state);
    DeviceFree(m->treeMem);
    delete m;
}

int32_t MetricsSetTree(hwpa_metrics* m, const hwpa_activity_tree* t)
{
    if (!m) return Fail(HWPA_ERR_INVALID_ARG, "null metrics");
    HWPA_HIP(hipDeviceSynchronize());
    DeviceFree(m->treeMem);
    m->treeMem = nullptr;
    m->tree = TreeDev{};
    if (!t || t->num_nodes == 0) return HWPA_OK;
    const int32_t n = t->num_nodes;
    if (n < 0 || n > kMaxTreeNodes || !t->children_left || !t->children_right || !t->feature || !t->threshold ||
        !t->value)
        return Fail(HWPA_ERR_INVALID_ARG, "activity tree: bad size or null array");
    for (int32_t i = 0; i < n; ++i) {
        const int f = t->feature[i];
        if (f >= kTreeFeatures) return Fail(HWPA_ERR_INVALID_ARG, "activity tree: feature index out of range");
        if (f >= 0 && (t->children_left[i] <= i || t->children_left[i] >= n || t->children_right[i] <= i ||
                       t->children_right[i] >= n))
            return Fail(HWPA_ERR_INVALID_ARG, "activity tree: a child must follow its parent");
    }
    // One allocation: int16 left, int16 right, int8 feature, int8 value, float threshold.
    const size_t oL = 0, oR = oL + 2 * (size_t)n, oF = oR + 2 * (size_t)n, oV = oF + (size_t)n;
    const size_t oT = (oV + (size_t)n + 15) / 16 * 16, total = oT + 4 * (size_t)n;
    std::vector<uint8_t> host(total, 0);
    std::memcpy(host.data() + oL, t->children_left, 2 * (size_t)n);
    std::memcpy(host.data() + oR, t->children_right, 2 * (size_t)n);
    std::memcpy(host.data() + oF, t->feature, (size_t)n);
    std::memcpy(host.data() + oV, t->value, (size_t)n);
    std::memcpy(host.data() + oT, t->threshold, 4 * (size_t)n);
    int32_t rc = DeviceAlloc(&m->treeMem, total, "activity tree");
    if (rc != HWPA_OK) return rc;
    if (hipMemcpy(m->treeMem, host.data(), total, hipMemcpyHostToDevice) != hipSuccess)
        return Fail(HWPA_ERR_HIP, "activity tree upload failed");
    uint8_t* base = (uint8_t*)m->treeMem;
    TreeDev d;
    d.nodes = n;
    d.maxHsw = t->max_acceptable_hsw_rate;
    d.hsw[0] = t->hsw_curve[0];
    d.hsw[1] = t->hsw_curve[1];
    d.left = (const int16_t*)(base + oL);
    d.right = (const int16_t*)(base + oR);
    d.feature = (const int8_t*)(base + oF);
    d.value = (const int8_t*)(base + oV);
    d.threshold = (const float*)(base + oT);
    m->tree = d;
    return HWPA_OK;
}

// HFMetricsFilterHost::Process (HFMetricsFilterHost.cpp:86-116): the block
// bookkeeping is host-side, so *produced is known when this returns.
int32_t MetricsProcess(hwpa_metrics* m, const hwpa_pulse* pulses, const uint32_t* counts, uint32_t maxCalls,
                       uint32_t frames, int32_t lastFrame, const hwpa_baseliner_stats* blStats, const float* score,
                       const hwpa_stat_acc* pdStats, const hwpa_lane_model* models, hwpa_basecalling_metrics* out,
                       int32_t* produced, hipStream_t s)
{
    if (!m || !pulses || !counts || !blStats || !score || !pdStats || !models || !out || maxCalls == 0 || frames == 0)
        return Fail(HWPA_ERR_INVALID_ARG, "metrics_process: bad argument");
    const int32_t reset = (m->framesSeen == 0) ? 1 : 0;
    hipLaunchKernelGGL(HfmChunkKernel, dim3(m->lanes), dim3(LS), 0, s, m->state, pulses, counts, maxCalls, frames,
                       blStats, score, pdStats, models, reset);
    int32_t rc = CheckLaunch("HfmChunkKernel");
    if (rc != HWPA_OK) return rc;
    m->framesSeen += frames;
    const bool end = (uint32_t)lastFrame % m->cfg.frames_per_hf_metric_block < frames;
    if (produced) *produced = end ? 1 : 0;
    hipLaunchKernelGGL(HfmFinalizeKernel, dim3(m->lanes), dim3(LS), 0, s, m->state, m->tree,
                       m->cfg.realtime_activity_labels, m->cfg.movie.frame_rate, out, end ? 1 : 0);
    if (end) m->framesSeen = 0;
    return CheckLaunch("HfmFinalizeKernel");
}

}  // namespace hwpa

extern "C" {

int32_t hwpa_metrics_create(const hwpa_config* cfg, uint32_t lanes, hwpa_metrics** out)
{
    if (!cfg || !out) return Fail(HWPA_ERR_INVALID_ARG, "null argument");
    const int32_t rc = hwpa_config_validate(cfg);
    if (rc != HWPA_OK) return rc;
    return MetricsCreate(*cfg, lanes, out);
}

int32_t hwpa_metrics_destroy(hwpa_metrics* m)
{
    MetricsDestroy(m);
    return HWPA_OK;
}

int32_t hwpa_metrics_set_activity_tree(hwpa_metrics* m, const hwpa_activity_tree* tree)
{
    return MetricsSetTree(m, tree);
}

int32_t hwpa_metrics_process(hwpa_metrics* m, const hwpa_pulse* d_pulses, const uint32_t* d_counts,
                             uint32_t max_calls, uint32_t frames, int32_t last_frame,
                             const hwpa_baseliner_stats* d_bl_stats, const float* d_viterbi_score,
                             const hwpa_stat_acc* d_pd_stats, const hwpa_lane_model* d_models,
                             hwpa_basecalling_metrics* d_out, int32_t* produced, void* stream)
{
    return MetricsProcess(m, d_pulses, d_counts, max_calls, frames, last_frame, d_bl_stats, d_viterbi_score,
                          d_pd_stats, d_models, d_out, produced, (hipStream_t)stream);
}

}  // extern "C"
