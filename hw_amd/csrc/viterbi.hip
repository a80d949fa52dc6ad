// viterbi.hip -- frame labeling: ROI hysteresis prefilter, 13-state subframe
// Viterbi with a 16-frame stitch look-back, traceback and latent update
// (FrameLabelerHost, traceAnalysis/FrameLabelerHost.cpp:53-612; scorer,
// traceAnalysis/SubframeScorer.h:143-318; ROI, traceAnalysis/RoiHost.h:75-248;
// transition matrix, traceAnalysis/SubframeScorer.cpp:35-223).
//
// Mapping: one wave64 per lane, one thread per ZMW, sequential over frames.
// The ROI is computed first (forward hysteresis, then a backward pass) into a
// 2-bit-per-frame LDS array; the recursion then runs over latent + chunk +
// stitch frames.  The transition matrix has a fixed sparsity (each row: one
// leading column plus one contiguous segment), so the whole predecessor
// choice of a frame packs into one 32-bit word: 4 bits for the baseline row
// and each up row (9 candidates), 1 bit for each full and down row (2
// candidates).  One coalesced 256-byte word row per frame goes to HBM and is
// read back once by the traceback.  Ties pick the later predecessor, as the
// host recursion's `>=` does.

#include "internal.h"

#include <cmath>

namespace hwpa {

struct TransVals {
    float bu[5][5];   // rows B, TU, GU, CU, AU: candidates {B, TU, GU, CU, AU}
    float bd[5];      // the same rows: the one value shared by candidates TD, GD, CD, AD
    float fd[8][2];   // rows T, G, C, A, TD, GD, CD, AD: candidates {full_i, up_i}
};

struct LabLaneState {
    hwpa_lane_model oldModel;
    int16_t labelBoundary[LS];
    uint16_t roiBoundary[LS];
    int16_t latentTrc[LS];
    int16_t prevLat[NL][LS];
};

struct LabelerDev {
    TransVals T;
    float roiUpper, roiLower;
    uint32_t lanes, maxFrames;
    LabLaneState* state = nullptr;
    uint32_t* tb = nullptr;  // [lane][maxFrames + 16 + NS][LS]: predecessor words, then the anchor row
};

// nhB = -1/(2 var_B), nhP = -1/(2 var_a): the reference's -0.5f * y * invVar * y
// equals (y * (-0.5f * invVar)) * y bit for bit (scaling by -0.5 is exact here).
struct ScorerZ {
    float bgMean, nhB, bgFixed, x0;
    float ffMean[NA], nhP[NA], ffFixed[NA], logPMean[NA], x1[NA];
    float dimMean, dimNorm, dimScale;
};

// BlockStateSubframeScorer::Setup / SubframeSetup (SubframeScorer.h:151-281)
__device__ void ScorerSetup(ScorerZ& s, const hwpa_lane_model& m, int z)
{
    const float log2pi = 1.8378770664f;
    const float pi = 3.1415926536f;
    const float sqrtHalfPi = hw_sqrtf(0.5f * pi);
    const float bMean = hw_h2f(m.baseline.means[z]);
    const float bVar = hw_h2f(m.baseline.vars[z]);
    const float bSigma = hw_sqrtf(bVar);
    s.x0 = bMean + bSigma * sqrtHalfPi;
    s.nhB = -0.5f * (1.0f / bVar);
#pragma unroll
    for (int a = 0; a < NA; ++a) {
        const float aMean = hw_h2f(m.analogs[a].means[z]);
        s.logPMean[a] = hw_cephes_logf(aMean - bMean);
        const float aVar = hw_h2f(m.analogs[a].vars[z]);
        const float aSigma = hw_sqrtf(aVar);
        s.x1[a] = aMean - aSigma * sqrtHalfPi;
        s.nhP[a] = -0.5f * (1.0f / aVar);
    }
    {
        const float aMean = hw_h2f(m.analogs[3].means[z]);
        const float aVar = hw_h2f(m.analogs[3].vars[z]);
        const float eMean = 0.5f * (bMean + aMean);
        float eVar = (1.0f / 6.0f) * (aMean - bMean);
        eVar += bVar + aVar;
        eVar *= 0.5f;
        s.dimMean = eMean;
        s.dimNorm = -0.5f * (log2pi + hw_cephes_logf(eVar));
        s.dimScale = 0.5f / eVar;
    }
    const float normConst = log2pi / 2.0f;
    s.bgFixed = -0.5f * hw_cephes_logf(bVar) - normConst;
    s.bgMean = bMean;
#pragma unroll
    for (int a = 0; a < NA; ++a) {
        s.ffFixed[a] = -0.5f * hw_cephes_logf(hw_h2f(m.analogs[a].vars[z])) - normConst;
        s.ffMean[a] = hw_h2f(m.analogs[a].means[z]);
    }
}

// BlockStateSubframeScorer::StateScores for all 13 states (SubframeScorer.h:182-228, 284-305)
__device__ __forceinline__ void Scores(const ScorerZ& s, float d, float* sc)
{
    {
        const float y = d - s.bgMean;
        sc[0] = (y * s.nhB) * y + s.bgFixed;
    }
    float ff[NA];
#pragma unroll
    for (int i = 0; i < NA; ++i) {
        const float y = d - s.ffMean[i];
        ff[i] = (y * s.nhP[i]) * y + s.ffFixed[i];
        sc[1 + i] = ff[i];
    }
    const bool lowExt = d < s.x0;
#pragma unroll
    for (int i = 0; i < NA; ++i) {
        const float cap = lowExt ? s.x0 : s.x1[i];
        const float fixedTerm = lowExt ? s.nhB : s.nhP[i];
        const bool ext = lowExt || (d > s.x1[i]);
        const float dc = d - cap;
        float v = (ext ? fixedTerm * (dc * dc) : 0.0f) - s.logPMean[i];
        if (i == NA - 1) {
            const float dm = d - s.dimMean;
            const float fb = s.dimNorm - (dm * dm) * s.dimScale;
            v = (s.x0 <= s.x1[i]) ? v : fb;
        }
        if (d > s.ffMean[i]) v = hw_minf(hw_minf(v, ff[i] - 1.0f), 0.0f);
        sc[5 + i] = v;   // up state
        sc[9 + i] = v;   // down state (same analog, same score)
    }
}

// One Viterbi recursion step (FrameLabelerHost.cpp:213-264): (score + transition) +
// previous log-likelihood per candidate, the later candidate winning ties.  In rows B
// and TU..AU the four down candidates share one transition value, so score + value is
// formed once per row (the same rounded sum the reference forms four times).
// pen = 0 inside the ROI, +inf outside (every state but the baseline is excluded).
__device__ __forceinline__ uint32_t Recur(const TransVals& T, const float* sc, float* ll, float pen)
{
    float acc[NS];
    uint32_t word = 0;
    // The five rows' maximum chains advance together, one candidate at a time, so
    // each compare has independent work behind it before its select.
    float mx[5], sr[5];
    uint32_t code[5];
#pragma unroll
    for (int q = 0; q < 5; ++q) {
        sr[q] = sc[(q == 0) ? 0 : 4 + q];
        mx[q] = (sr[q] + T.bu[q][0]) + ll[0];
        code[q] = 0;
    }
#pragma unroll
    for (int p = 0; p < 8; ++p) {
        if (p == 4) {
#pragma unroll
            for (int q = 0; q < 5; ++q) sr[q] = sr[q] + T.bd[q];
        }
#pragma unroll
        for (int q = 0; q < 5; ++q) {
            const float v = (p < 4) ? (sr[q] + T.bu[q][1 + p]) + ll[5 + p] : sr[q] + ll[5 + p];
            const bool c = v >= mx[q];
            mx[q] = c ? v : mx[q];
            code[q] = c ? (uint32_t)(1 + p) : code[q];
        }
    }
#pragma unroll
    for (int q = 0; q < 5; ++q) {
        const int r = (q == 0) ? 0 : 4 + q;
        acc[r] = (r == 0) ? mx[q] : (mx[q] - pen);
        word |= code[q] << (4 * q);
    }
    // full rows 1..4 and down rows 9..12: candidates {full_i, up_i}
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int i = q & 3;
        const int r = (q < 4) ? 1 + i : 9 + i;
        const float s = sc[r];
        const float a = (s + T.fd[q][0]) + ll[1 + i];
        const float b = (s + T.fd[q][1]) + ll[5 + i];
        const bool c = b >= a;
        const float m = c ? b : a;
        acc[r] = m - pen;
        word |= (c ? 1u : 0u) << (20 + q);
    }
#pragma unroll
    for (int r = 0; r < NS; ++r) ll[r] = acc[r];
    return word;
}

__device__ __forceinline__ int Pred(uint32_t w, int state)
{
    if (state == 0 || (state >= 5 && state <= 8)) {
        const int q = (state == 0) ? 0 : state - 4;
        const int code = (int)((w >> (4 * q)) & 15u);
        return code == 0 ? 0 : code + 4;
    }
    const int i = (state <= 4) ? state - 1 : state - 9;
    const int q = (state <= 4) ? i : 4 + i;
    return ((w >> (20 + q)) & 1u) ? 5 + i : 1 + i;
}

// Pred for a state that differs between ZMWs (the traceback): both field lookups and a
// select, no divergent branch.  Shift amounts stay below 32 for every state.
__device__ __forceinline__ int PredSel(uint32_t w, int state)
{
    const uint32_t qb = (state == 0) ? 0u : ((uint32_t)(state - 4) & 7u);
    const int code = (int)((w >> (4u * qb)) & 15u);
    const int pb = code == 0 ? 0 : code + 4;
    const int i = (state <= 4 ? state - 1 : state - 9) & 3;
    const uint32_t q = (uint32_t)((state <= 4) ? i : 4 + i);
    const int pf = ((w >> (20u + q)) & 1u) ? 5 + i : 1 + i;
    const bool bu = state == 0 || (state >= 5 && state <= 8);
    return bu ? pb : pf;
}

// The ROI penalty of the frame whose final ROI bit is bit `pos` of `w`: 0 or +inf.
__device__ __forceinline__ float RoiPen(uint32_t w, uint32_t pos)
{
    const uint32_t m = (uint32_t)__builtin_amdgcn_sbfe((int32_t)w, pos, 1);   // 0 or all ones
    return hw_u2f(~m & 0x7f800000u);
}

// PB = frames per prefetched batch (8 or 16) in the recursion and the traceback.
// MINW = waves per SIMD the register allocation must allow (1: unconstrained).
template <int PB, int MINW>
__global__ __launch_bounds__(64, MINW) void LabelKernel(const int16_t* __restrict__ traces, int frames,
                                                  const hwpa_lane_model* __restrict__ models, LabLaneState* state,
                                                  uint32_t* __restrict__ tbAll, int tbFrames,
                                                  int16_t* __restrict__ labels, int16_t* __restrict__ latentOut,
                                                  float* __restrict__ score, TransVals T, float roiUpper,
                                                  float roiLower)
{
    static_assert(PB % 8 == 0, "recursion batches cover whole 8-frame groups of an ROI word");
    extern __shared__ uint32_t roiWords[];  // [(nroi + 15) / 16][LS], 2 bits per frame
    const int lane = blockIdx.x;
    const int z = threadIdx.x;
    LabLaneState& st = state[lane];
    const hwpa_lane_model& model = models[lane];
    const int16_t* tr = traces + (size_t)lane * frames * LS + z;
    uint32_t* tb = tbAll + (size_t)lane * tbFrames * LS + z;
    const int nroi = frames + HWPA_VITERBI_LOOKBACK + HWPA_ROI_STITCH;
    const int nlab = frames + HWPA_VITERBI_LOOKBACK;
    const int nwords = (nroi + 15) / 16;

    // Latent trace out = the 24 frames preceding this chunk.
    for (int i = 0; i < NL; ++i) latentOut[((size_t)lane * NL + i) * LS + z] = st.prevLat[i][z];

    const float inv1 = 1.0f / hw_sqrtf(hw_h2f(st.oldModel.baseline.vars[z]));
    const float inv2 = 1.0f / hw_sqrtf(hw_h2f(model.baseline.vars[z]));
    // Frame k of the latent ++ chunk sequence (k < NL: the carried latent frames).
    auto rawAt = [&](int k) -> int { return (k < NL) ? (int)st.prevLat[k][z] : (int)tr[(size_t)(k - NL) * LS]; };

    // ROI forward pass (RoiHost.h:137-193) over k < nroi, reading frame k + 2, one
    // 16-frame word at a time with the next word's frames in flight; indices past
    // the sequence re-read its last frame (their values are never used).
    {
        const float kS2 = 1.41421353816986083984375f;   // sqrt(2.0f)
        const float kS3 = 1.73205077648162841796875f;   // sqrt(3.0f)
        const float kR2 = 1.0f / kS2, kR3 = 1.0f / kS3;
        float p0 = (float)st.latentTrc[z] * inv1, p1 = (float)rawAt(0) * inv1, p2 = (float)rawAt(1) * inv1;
        uint32_t prevRoi = st.roiBoundary[z];
        const int limit = nroi + 2;
        auto load16 = [&](int base, int (&dst)[16]) {
#pragma unroll
            for (int u = 0; u < 16; ++u) dst[u] = rawAt(min(base + u, limit - 1));
        };
        auto frame = [&](int k, int raw) -> uint32_t {
            const float nv = (float)raw * ((k + 2 < NL) ? inv1 : inv2);
            const float a2 = p0 + p1;
            const float a3 = a2 + p2;
            const float a4 = a3 + nv;
            float mval = hw_maxf(p1, hw_div_by_const(a2, kS2, kR2));
            mval = hw_maxf(mval, hw_div_by_const(a3, kS3, kR3));
            mval = hw_maxf(mval, a4 * 0.5f);
            const float thresh = (prevRoi & 1u) ? roiLower : roiUpper;
            const uint32_t v = (mval >= thresh ? 1u : 0u) | (mval >= roiLower ? 2u : 0u);
            prevRoi = v;
            p0 = p1;
            p1 = p2;
            p2 = nv;
            return v;
        };
        int cur[16], nxt[16];
        load16(2, cur);
        const int full = nroi / 16;
        for (int wi = 0; wi < full; ++wi) {
            load16(16 * wi + 18, nxt);
            uint32_t word = 0;
#pragma unroll
            for (int u = 0; u < 16; ++u) word |= frame(16 * wi + u, cur[u]) << (2 * u);
            roiWords[wi * LS + z] = word;
#pragma unroll
            for (int u = 0; u < 16; ++u) cur[u] = nxt[u];
        }
        if (full < nwords) {
            uint32_t word = 0;
#pragma unroll
            for (int u = 0; u < 16; ++u)
                if (16 * full + u < nroi) word |= frame(16 * full + u, cur[u]) << (2 * u);
            roiWords[full * LS + z] = word;
        }
    }
    // ROI backward pass (RoiHost.h:195-243), a word at a time: the final ROI bit of
    // each frame into bit 0 of its field (fields past the last frame stay zero).
    {
        uint32_t carry = 0;
        for (int wi = nwords - 1; wi >= 0; --wi)
            roiWords[wi * LS + z] = hw_roi_backward_word(roiWords[wi * LS + z], &carry);
    }

    // Forward recursion over latent, chunk and stitch frames.
    float ll[NS];
    const int lb = st.labelBoundary[z];
#pragma unroll
    for (int s = 0; s < NS; ++s) ll[s] = (lb == s) ? 0.0f : hw_u2f(0xff800000u);
    ScorerZ sc;
    auto step = [&](int f, int raw, uint32_t rw) {
        float s[NS];
        Scores(sc, (float)raw, s);
        tb[(size_t)f * LS] = Recur(T, s, ll, RoiPen(rw, 2u * (uint32_t)(f & 15)));
    };
    // The latent frames are scored with the previous chunk's model, the chunk
    // and stitch frames with the current one.
    ScorerSetup(sc, st.oldModel, z);
    for (int f = 0; f < NL; ++f) step(f, (int)st.prevLat[f][z], roiWords[(f >> 4) * LS + z]);
    // Log-likelihoods at the chunk's last frame, read back once the anchor state is known:
    // kept in the last NS rows of the lane's traceback buffer (13 registers fewer through
    // the stitch frames, no LDS).
    uint32_t* anchorLL = tb + (size_t)(tbFrames - NS) * LS;
    auto storeAnchor = [&]() {
#pragma unroll
        for (int r = 0; r < NS; ++r) anchorLL[(size_t)r * LS] = hw_f2u(ll[r]);
    };
    ScorerSetup(sc, model, z);
    {
        // Chunk frames f in [NL, frames) in batches of PB (f0 a multiple of 8, so each
        // group of 8 frames reads one ROI word) with the next batch in flight, then
        // the remainder; chunk sample j = f - NL.
        const int last = nlab - NL - 1;   // last chunk sample the recursion reads
        auto loadBatch = [&](int j0, int (&dst)[PB]) {
#pragma unroll
            for (int u = 0; u < PB; ++u) dst[u] = (int)tr[(size_t)min(j0 + u, last) * LS];
        };
        const int nb = (frames - NL) / PB;
        int cur[PB], nxt[PB];
        loadBatch(0, cur);
        for (int b = 0; b < nb; ++b) {
            const int f0 = NL + b * PB;
            loadBatch((b + 1) * PB, nxt);
#pragma unroll
            for (int g = 0; g < PB / 8; ++g) {
                const int fg = f0 + 8 * g;
                const uint32_t rw = roiWords[(fg >> 4) * LS + z];
#pragma unroll
                for (int u = 0; u < 8; ++u) step(fg + u, cur[8 * g + u], rw);
            }
#pragma unroll
            for (int u = 0; u < PB; ++u) cur[u] = nxt[u];
        }
        // The remainder of the chunk frames, the anchor, then the stitch frames f in
        // [frames, nlab) (chunk samples frames - NL .. last), one frame at a time with
        // the next sample in flight.
        const int f1 = NL + nb * PB;
        int raw = (int)tr[(size_t)(f1 - NL) * LS];
#pragma unroll 1
        for (int f = f1; f < nlab; ++f) {
            const int nextRaw = (int)tr[(size_t)min(f + 1 - NL, last) * LS];
            if (f == frames) storeAnchor();
            step(f, raw, roiWords[(f >> 4) * LS + z]);
            raw = nextRaw;
        }
    }

    // Normalize and propagate end-state probabilities back 16 frames (FrameLabelerHost.cpp:185-205, 388-417)
    float prob[NS];
    {
        float mx = ll[0];
#pragma unroll
        for (int i = 1; i < NS; ++i) mx = hw_maxf(ll[i], mx);
        float sum = 0.0f;
#pragma unroll
        for (int i = 0; i < NS; ++i) {
            prob[i] = hw_cephes_expf(ll[i] - mx);
            sum += prob[i];
        }
        const float inv = 1 / sum;
#pragma unroll
        for (int i = 0; i < NS; ++i) prob[i] *= inv;
    }
    for (int i = nlab - 1; i > frames - 1; --i) {
        const uint32_t w = tb[(size_t)i * LS];
        float np[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) np[s] = 0.0f;
#pragma unroll
        for (int from = 0; from < NS; ++from) {
            const int ps = Pred(w, from);
#pragma unroll
            for (int to = 0; to < NS; ++to) np[to] += (to == ps) ? prob[from] : 0.0f;
        }
#pragma unroll
        for (int s = 0; s < NS; ++s) prob[s] = np[s];
    }
    float maxProb = prob[0];
    int anchorState = 0;
#pragma unroll
    for (int i = 1; i < NS; ++i) {
        const bool cond = maxProb > prob[i];
        maxProb = cond ? maxProb : prob[i];
        anchorState = cond ? anchorState : i;
    }
    const float sv = hw_u2f(anchorLL[(size_t)anchorState * LS]);
    score[(size_t)lane * LS + z] = sv;

    // Traceback from the anchor, PB predecessor words per batch with the next batch in
    // flight, then the remainder.
    int16_t* lo = labels + (size_t)lane * frames * LS + z;
    int stt = anchorState;
    {
        auto loadTb = [&](int top, uint32_t (&dst)[PB]) {
#pragma unroll
            for (int u = 0; u < PB; ++u) dst[u] = tb[(size_t)max(top - u, 0) * LS];
        };
        const int nb = frames / PB;
        uint32_t cur[PB], nxt[PB];
        loadTb(frames - 1, cur);
        for (int b = 0; b < nb; ++b) {
            const int top = frames - 1 - b * PB;
            loadTb(top - PB, nxt);
#pragma unroll
            for (int u = 0; u < PB; ++u) {
                lo[(size_t)(top - u) * LS] = (int16_t)stt;
                stt = PredSel(cur[u], stt);
            }
#pragma unroll
            for (int u = 0; u < PB; ++u) cur[u] = nxt[u];
        }
#pragma unroll 1
        for (int f = frames - 1 - nb * PB; f >= 0; --f) {
            lo[(size_t)f * LS] = (int16_t)stt;
            stt = PredSel(tb[(size_t)f * LS], stt);
        }
    }

    // Latent update
    const int anchor = frames - NL;
    const int fl = frames - 1;
    st.roiBoundary[z] = (uint16_t)((roiWords[(fl >> 4) * LS + z] >> (2 * (fl & 15))) & 1u);
    st.labelBoundary[z] = (int16_t)anchorState;
    st.latentTrc[z] = tr[(size_t)(anchor - 1) * LS];
    for (int i = 0; i < NL; ++i) st.prevLat[i][z] = tr[(size_t)(anchor + i) * LS];
    // Model copy (each thread copies its ZMW column)
    st.oldModel.baseline.means[z] = model.baseline.means[z];
    st.oldModel.baseline.vars[z] = model.baseline.vars[z];
    st.oldModel.baseline.weights[z] = model.baseline.weights[z];
#pragma unroll
    for (int a = 0; a < NA; ++a) {
        st.oldModel.analogs[a].means[z] = model.analogs[a].means[z];
        st.oldModel.analogs[a].vars[z] = model.analogs[a].vars[z];
        st.oldModel.analogs[a].weights[z] = model.analogs[a].weights[z];
    }
    st.oldModel.confidence[z] = model.confidence[z];
}

__global__ __launch_bounds__(64) void LabelInitKernel(LabLaneState* state)
{
    LabLaneState& st = state[blockIdx.x];
    const int z = threadIdx.x;
    // LatentViterbi ctor (FrameLabelerHost.cpp:58-81)
    st.oldModel.baseline.means[z] = hw_f2h(0.0f);
    st.oldModel.baseline.vars[z] = hw_f2h(50.0f);
    st.oldModel.baseline.weights[z] = 0;
    for (int a = 0; a < NA; ++a) {
        st.oldModel.analogs[a].means[z] = hw_f2h(1000.0f);
        st.oldModel.analogs[a].vars[z] = hw_f2h(10.0f);
        st.oldModel.analogs[a].weights[z] = 0;
    }
    st.oldModel.confidence[z] = 0;
    st.labelBoundary[z] = 0;
    st.roiBoundary[z] = 0;
    st.latentTrc[z] = 0;
    for (int i = 0; i < NL; ++i) st.prevLat[i][z] = 0;
}

// TransitionMatrix<float> ctor (SubframeScorer.cpp:35-223), computed on the host.
void BuildTransition(const hwpa_config& cfg, hwpa_transition& T)
{
    hwpa_analog analogs[NA];
    for (int i = 0; i < NA; ++i) analogs[i] = cfg.movie.analogs[i];
    float data[NS][NS];
    for (int i = 0; i < NS; ++i)
        for (int j = 0; j < NS; ++j) data[i][j] = 0.0f;
    for (int i = 0; i < NA; ++i) {
        const float delta = 0.01f;
        // The first check tests the IPD ratio but adjusts the pulse-width ratio,
        // as the reference does.
        if (std::abs(analogs[i].ipd2_slow_step_ratio - 1.0f) < delta)
            analogs[i].pw2_slow_step_ratio = (analogs[i].pw2_slow_step_ratio < 1.0f) ? 1.0f - delta : 1.0f + delta;
        if (std::abs(analogs[i].ipd2_slow_step_ratio - 1.0f) < delta)
            analogs[i].ipd2_slow_step_ratio = (analogs[i].ipd2_slow_step_ratio < 1.0f) ? 1.0f - delta : 1.0f + delta;
    }
    float meanIpd = 0.0f;
    for (int i = 0; i < NA; ++i) meanIpd += analogs[i].inter_pulse_distance;
    meanIpd = meanIpd / NA;
    const double frameRate = cfg.movie.frame_rate;
    // Probabilities of 1-, 2- and 3-frame events for a two-step kinetic model
    // (double precision, as in the host computation).
    auto lenProbs = [&](float widthSeconds, float ratio, float* ret) {
        const double widthFrames = frameRate * widthSeconds;
        ratio = std::max(ratio, 1e-6f);
        const double k1 = (1.0f + ratio) / widthFrames;
        const double k2 = (1.0f + 1.0f / ratio) / widthFrames;
        const double dk = k1 - k2;
        const double c1 = 1.0f - std::exp(-k1);
        const double c2 = 1.0f - std::exp(-k2);
        ret[0] = (float)(1 - k1 * c2 / k2 / dk + k2 * c1 / k1 / dk);
        ret[1] = (float)(k1 * (c2 * c2) / k2 / dk - k2 * (c1 * c1) / k1 / dk);
        ret[2] = (float)(k1 * (c2 * c2) * (1 - c2) / k2 / dk - k2 * (c1 * c1) * (1 - c1) / k1 / dk);
    };
    float ipdLen[NA][3];
    for (int i = 0; i < NA; ++i) lenProbs(analogs[i].inter_pulse_distance, analogs[i].ipd2_slow_step_ratio, ipdLen[i]);
    float alpha[NA];
    for (int i = 0; i < NA; ++i) alpha[i] = cfg.viterbi_alpha * 0.25f * (ipdLen[i][0] + ipdLen[i][1]);
    float alphaSum = 0.0f;
    for (int i = 0; i < NA; ++i) alphaSum += alpha[i];
    float meanLen[3];
    lenProbs(meanIpd, 0.0f, meanLen);
    const float meanIpd3m = 1 - meanLen[0] - meanLen[1];
    const float pulseStartProb = meanLen[2] / meanIpd3m;
    data[0][0] = 1.0f - pulseStartProb;
    for (int i = 0; i < NA; ++i) {
        const int full = i + 1, up = i + 1 + NA, down = i + 1 + 2 * NA;
        float al[3];
        lenProbs(analogs[i].pulse_width, analogs[i].pw2_slow_step_ratio, al);
        const float gamma = cfg.viterbi_gamma * al[0] / (1 + alphaSum);
        const float beta = cfg.viterbi_beta * al[1];
        data[up][0] = 0.25f * pulseStartProb;
        data[0][up] = gamma;
        data[down][up] = beta;
        data[full][up] = 1 - gamma - gamma * alphaSum - beta;
        const float prob3ms = 1.0f - al[0] - al[1];
        data[down][full] = al[2] / prob3ms;
        data[full][full] = 1.0f - data[down][full];
        data[0][down] = 1 - alphaSum;
        for (int j = 0; j < NA; ++j) {
            const int upPrime = j + 1 + NA;
            data[upPrime][down] = alpha[j];
            data[upPrime][up] = gamma * alpha[j];
        }
    }
    for (int r = 0; r < NS; ++r) {
        if (r == 0 || (r >= 5 && r <= 8)) {
            T.first_col[r] = 0;
            T.seg_first[r] = 5;
            T.seg_last[r] = 13;
        } else {
            const int a = (r <= 4) ? r - 1 : r - 9;
            T.first_col[r] = a + 1;
            T.seg_first[r] = a + 5;
            T.seg_last[r] = a + 6;
        }
    }
    // Entries with probability <= 0 keep the storage's initial 0 (SubframeScorer.cpp:213-220).
    for (int i = 0; i < NS; ++i)
        for (int j = 0; j < NS; ++j) T.logp[i][j] = (data[i][j] > 0.0f) ? std::log(data[i][j]) : 0.0f;
}

// Packs the transition matrix into the kernel's candidate order.  Fails when the
// four down candidates of a row B / TU..AU do not share one value (the matrix the
// reference builds always has that form: SubframeScorer.cpp:173-200).
static bool PackTransition(const hwpa_transition& T, TransVals& v)
{
    v = TransVals{};
    const int buRows[5] = {0, 5, 6, 7, 8};
    for (int q = 0; q < 5; ++q) {
        const int r = buRows[q];
        v.bu[q][0] = T.logp[r][0];
        for (int p = 0; p < 4; ++p) v.bu[q][1 + p] = T.logp[r][5 + p];
        v.bd[q] = T.logp[r][9];
        for (int p = 10; p < 13; ++p)
            if (hw_f2u(T.logp[r][p]) != hw_f2u(v.bd[q])) return false;
    }
    for (int q = 0; q < 8; ++q) {
        const int i = q & 3;
        const int r = (q < 4) ? 1 + i : 9 + i;
        v.fd[q][0] = T.logp[r][1 + i];
        v.fd[q][1] = T.logp[r][5 + i];
    }
    return true;
}

int32_t LabelerCreate(const hwpa_config& cfg, uint32_t lanes, uint32_t max_frames, LabelerDev** out)
{
    if (max_frames < (uint32_t)NL) return Fail(HWPA_ERR_INVALID_ARG, "labeler: max_frames must be >= 24");
    hwpa_transition T;
    BuildTransition(cfg, T);
    TransVals packed;
    if (!PackTransition(T, packed))
        return Fail(HWPA_ERR_INVALID_ARG, "labeler: transition rows do not share their down-state values");
    auto* l = new LabelerDev;
    l->T = packed;
    l->roiUpper = cfg.roi_upper_threshold;
    l->roiLower = cfg.roi_lower_threshold;
    l->lanes = lanes;
    l->maxFrames = max_frames;
    int32_t rc = DeviceAlloc((void**)&l->state, sizeof(LabLaneState) * lanes, "labeler state");
    if (rc == HWPA_OK)
        rc = DeviceAlloc((void**)&l->tb, (size_t)lanes * (max_frames + HWPA_VITERBI_LOOKBACK + NS) * LS * sizeof(uint32_t),
                         "labeler traceback");
    if (rc == HWPA_OK) {
        hipLaunchKernelGGL(LabelInitKernel, dim3(lanes), dim3(LS), 0, nullptr, l->state);
        rc = CheckLaunch("LabelInitKernel");
    }
    if (rc == HWPA_OK && hipDeviceSynchronize() != hipSuccess) rc = Fail(HWPA_ERR_HIP, "labeler init failed");
    if (rc != HWPA_OK) {
        LabelerDestroy(l);
        return rc;
    }
    *out = l;
    return HWPA_OK;
}

void LabelerDestroy(LabelerDev* l)
{
    if (!l) return;
    DeviceFree(l->state);
    DeviceFree(l->tb);
    delete l;
}

int32_t LabelerRun(LabelerDev* l, const int16_t* traces, uint32_t frames, const hwpa_lane_model* models,
                   int16_t* labels, int16_t* latent_out, float* score, hipStream_t s)
{
    if (!l || !traces || !models || !labels || !latent_out || !score)
        return Fail(HWPA_ERR_INVALID_ARG, "labeler_run: null argument");
    if (frames <= (uint32_t)NL || frames > l->maxFrames)
        return Fail(HWPA_ERR_INVALID_ARG, "labeler_run: frames outside (24, max_frames]");
    const int nroi = (int)frames + HWPA_VITERBI_LOOKBACK + HWPA_ROI_STITCH;
    const size_t lds = (size_t)((nroi + 15) / 16) * LS * sizeof(uint32_t);
    // Variants: frames per prefetched batch x registers held to 4 waves per SIMD.
#define HWPA_LABEL_LAUNCH(PB, W)                                                                                   \
    hipLaunchKernelGGL((LabelKernel<PB, W>), dim3(l->lanes), dim3(LS), lds, s, traces, (int)frames, models,        \
                       l->state, l->tb, (int)(l->maxFrames + HWPA_VITERBI_LOOKBACK + NS), labels, latent_out,      \
                       score, l->T, l->roiUpper, l->roiLower)
    switch (TuneGet(kTuneLabel, l->lanes, frames)) {
    case 1: HWPA_LABEL_LAUNCH(8, 1); break;
    case 2: HWPA_LABEL_LAUNCH(8, 4); break;
    case 3: HWPA_LABEL_LAUNCH(16, 4); break;
    default: HWPA_LABEL_LAUNCH(16, 1); break;
    }
#undef HWPA_LABEL_LAUNCH
    return CheckLaunch("LabelKernel");
}

}  // namespace hwpa

extern "C" int32_t hwpa_transition_matrix(const hwpa_config* cfg, hwpa_transition* out)
{
    if (!cfg || !out) return hwpa::Fail(HWPA_ERR_INVALID_ARG, "null argument");
    hwpa::BuildTransition(*cfg, *out);
    return HWPA_OK;
}
