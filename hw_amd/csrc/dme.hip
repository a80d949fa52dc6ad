// dme.hip -- detection-model estimation: model initialisation from baseline
// statistics and the grouped-data EM fit of the 5-component mixture
// (DmeEmHost, traceAnalysis/DmeEmHost.cpp:131-925; DetectionModelHost
// update/export, traceAnalysis/DetectionModelHost.cpp:69-285).
//
// Mapping: one wave64 per lane, one thread per ZMW.  The only coupling
// between ZMWs is the EM early exit, which the reference takes when every ZMW
// of the lane has converged or failed (DmeEmHost.cpp:485); here that is one
// wave vote per iteration.  Each EM iteration is one pass over the 301 bin
// boundaries in registers: the M-step's baseline-variance sum is carried in
// the same pass (no per-boundary scratch).  All arithmetic keeps the CPU
// restatement's operation order so results match it bit for bit.

#include "internal.h"

#include <cfloat>
#include <mutex>

namespace hwpa {

struct DmeDevCfg {
    float blVarMin, blVarMax, blVarNominal, signalScaler, shotVarCoeff;
    float refSnr;
    float relAmp[NA], cv[NA];
    float mixThresh0, mixThresh1, scaleSnrConfTol, pulseAmpReg, snrDrop, snrThresh0, snrThresh1, successThresh;
    float hl0, hl1;
    float log2pi;
    int32_t iterLimit;
    int32_t iterToLimit;
    int32_t updateMethod;
};

static DmeDevCfg MakeDevCfg(const hwpa_config& c)
{
    DmeDevCfg d{};
    const DmeConsts k = MakeDmeConsts(c);
    d.blVarMin = k.blVarMin;
    d.blVarMax = k.blVarMax;
    d.blVarNominal = k.blVarNominal;
    d.signalScaler = k.signalScaler;
    d.shotVarCoeff = k.shotVarCoeff;
    d.refSnr = c.movie.ref_snr;
    for (int a = 0; a < NA; ++a) {
        d.relAmp[a] = c.movie.analogs[a].rel_amplitude;
        d.cv[a] = c.movie.analogs[a].excess_noise_cv;
    }
    d.mixThresh0 = c.analog_mix_fraction_threshold[0];
    d.mixThresh1 = c.analog_mix_fraction_threshold[1];
    d.scaleSnrConfTol = c.scale_snr_conf_tol;
    d.pulseAmpReg = c.pulse_amp_regularization;
    d.snrDrop = c.snr_drop_thresh;
    d.snrThresh0 = c.min_analog_snr_thresh0;
    d.snrThresh1 = c.min_analog_snr_thresh1;
    d.successThresh = c.success_confidence_thresh;
    d.hl0 = c.confidence_half_life[0];
    d.hl1 = c.confidence_half_life[1];
    d.log2pi = std::log(2.0f * 3.1415926536f);   // libm, scalar (DmeEmHost.cpp:380)
    d.iterLimit = (int32_t)c.em_iteration_limit;
    d.iterToLimit = c.iterate_to_limit;
    d.updateMethod = (int32_t)c.model_update_method;
    return d;
}

struct Mode {
    float w, mean, var;
};
struct Model {
    Mode bl;
    Mode an[NA];
    float conf;
};

__device__ __forceinline__ float Covar(const DmeDevCfg& d, float cv2, float mean, float blVar)
{
    float r = blVar;
    r += d.shotVarCoeff * mean;
    r += cv2 * (mean * mean);
    return r;
}

__device__ __forceinline__ float Variance(float m0, float m1, float m2)
{
    float var = m1 * m1 / m0;
    var = (m2 - var) / (m0 - 1.0f);
    var = hw_maxf(var, 0.0f);
    return (m0 > 1.0f) ? var : hw_u2f(0x7fc00000u);
}

__device__ Model LoadModel(const hwpa_lane_model& m, int z)
{
    Model r;
    r.bl.w = hw_h2f(m.baseline.weights[z]);
    r.bl.mean = hw_h2f(m.baseline.means[z]);
    r.bl.var = hw_h2f(m.baseline.vars[z]);
#pragma unroll
    for (int a = 0; a < NA; ++a) {
        r.an[a].w = hw_h2f(m.analogs[a].weights[z]);
        r.an[a].mean = hw_h2f(m.analogs[a].means[z]);
        r.an[a].var = hw_h2f(m.analogs[a].vars[z]);
    }
    r.conf = hw_h2f(m.confidence[z]);
    return r;
}

__device__ void StoreModel(const Model& r, hwpa_lane_model& m, int z)
{
    m.baseline.weights[z] = hw_f2h(r.bl.w);
    m.baseline.means[z] = hw_f2h(r.bl.mean);
    m.baseline.vars[z] = hw_f2h(hw_minf(r.bl.var, 65000.0f));
#pragma unroll
    for (int a = 0; a < NA; ++a) {
        m.analogs[a].weights[z] = hw_f2h(r.an[a].w);
        m.analogs[a].means[z] = hw_f2h(r.an[a].mean);
        m.analogs[a].vars[z] = hw_f2h(hw_minf(r.an[a].var, 65000.0f));
    }
    m.confidence[z] = hw_f2h(r.conf);
}

// DetectionModelHost::Update0/1/2 + Update(other) (DetectionModelHost.cpp:69-205)
__device__ void UpdateConf(const DmeDevCfg& d, Model& t, const Model& o)
{
    const float confSum = t.conf + o.conf;
    const float f = (confSum > 0.0f) ? o.conf / confSum : 0.0f;
    const float a = 1.0f - f;
    const float b = f;
    if (d.updateMethod == 0) {
        t.bl.mean = a * t.bl.mean + b * o.bl.mean;
        t.bl.var = a * t.bl.var + b * o.bl.var;
#pragma unroll
        for (int i = 0; i < NA; ++i) {
            t.an[i].mean = a * t.an[i].mean + b * o.an[i].mean;
            t.an[i].var = a * t.an[i].var + b * o.an[i].var;
        }
    } else {
        const float prevBlVar = t.bl.var;
        const float bw = a * t.bl.w + b * o.bl.w;
        t.bl.w = bw;
        t.bl.mean = a * t.bl.mean + b * o.bl.mean;
        const float bv = hw_dpowf(t.bl.var, a) * hw_dpowf(o.bl.var, b);
        t.bl.var = bv;
        const float aw = 0.25f * (1.0f - bw);
#pragma unroll
        for (int i = 0; i < NA; ++i) {
            t.an[i].w = aw;
            float cv2;
            if (d.updateMethod == 2) {
                float tx = t.an[i].var - prevBlVar;
                tx -= t.an[i].mean * d.shotVarCoeff;
                tx /= (t.an[i].mean * t.an[i].mean);
                float ox = o.an[i].var - o.bl.var;
                ox -= o.an[i].mean * d.shotVarCoeff;
                ox /= (o.an[i].mean * o.an[i].mean);
                cv2 = a * tx + b * ox;
            } else {
                cv2 = d.cv[i] * d.cv[i];
            }
            const float am = hw_dpowf(t.an[i].mean, a) * hw_dpowf(o.an[i].mean, b);
            t.an[i].mean = am;
            t.an[i].var = Covar(d, cv2, am, bv);
        }
    }
    t.conf = confSum;
}

// DmeEmHost::InitLaneDetModel (DmeEmHost.cpp:871-925)
__global__ __launch_bounds__(64) void DmeInitKernel(const hwpa_baseliner_stats* stats, hwpa_lane_model* models,
                                                    DmeDevCfg d)
{
    const hwpa_baseliner_stats& s = stats[blockIdx.x];
    hwpa_lane_model& m = models[blockIdx.x];
    const int z = threadIdx.x;
    const float m0 = s.baseline.m0[z], m1 = s.baseline.m1[z], m2 = s.baseline.m2[z];
    const float blWeight = m0 / s.full.basic.m0[z];
    float blMean = m1 / m0 + s.baseline.offset[z];
    float blVar = Variance(m0, m1, m2);
    if (hw_isnan(blMean)) blMean = 0.0f;
    if (hw_isnan(blVar)) blVar = 100.0f;
    blVar = hw_clampf(blVar, d.blVarMin, d.blVarMax);
    m.confidence[z] = hw_f2h(0.1f);
    m.baseline.weights[z] = hw_f2h(blWeight);
    m.baseline.means[z] = hw_f2h(blMean);
    m.baseline.vars[z] = hw_f2h(blVar);
    const float amw = (1.0f - blWeight) / (float)NA;
    const float refSignal = d.refSnr * hw_sqrtf(blVar);
#pragma unroll
    for (int a = 0; a < NA; ++a) {
        m.analogs[a].weights[z] = hw_f2h(amw);
        const float aMean = hw_maxf(0.0f, blMean + d.relAmp[a] * refSignal);
        m.analogs[a].means[z] = hw_f2h(aMean);
        m.analogs[a].vars[z] = hw_f2h(Covar(d, d.cv[a] * d.cv[a], aMean, blVar));
    }
}

// Histogram view of one ZMW (UHistogramSimd(const LaneHistogram&), UHistogramSimd.cpp:100-118).
struct HistView {
    const hwpa_lane_hist* h;
    int z;
    float lb, bs;
    __device__ float Start(int b) const { return lb + (float)b * bs; }
    // Index NB is the always-empty overflow bin of UHistogramSimd (UHistogramSimd.cpp:66).
    __device__ uint32_t Cnt(int b) const { return (b < NB) ? (uint32_t)h->bin_count[b][z] : 0u; }
};

__device__ uint32_t InRange(const HistView& h, int* lastNonzero = nullptr)
{
    uint32_t s = 0;
    int last = -1;
    for (int i = 0; i < NB; ++i) {
        const uint32_t c = h.Cnt(i);
        s = (s + c) & 0xffffu;
        last = c ? i : last;
    }
    if (lastNonzero) *lastNonzero = last;
    return s;
}

// UHistogramSimd::CumulativeCount (UHistogramSimd.cpp:199-238)
__device__ float CumulativeCount(const HistView& h, uint32_t lowOut, uint32_t inRange, float x)
{
    const bool mNan = hw_isnan(x);
    const bool mLow = x < h.Start(0);
    const bool mHigh = x >= h.Start(NB);
    const bool mIn = !(mNan || mLow || mHigh);
    int xbin = hw_floor_i32((x - h.Start(0)) / h.bs);
    if (!mIn) xbin = 0;
    const float xrem = x - h.Start(xbin);
    const float xbinCount = (float)h.Cnt(xbin);
    float cc = (float)lowOut;
    uint32_t below = 0;
    for (int i = 0; i < xbin; ++i) below = (below + h.Cnt(i)) & 0xffffu;
    cc += (float)below;
    cc += xbinCount * xrem / h.bs;
    if (mLow) cc = (float)lowOut;
    if (mHigh) cc = (float)((lowOut + inRange) & 0xffffu);
    if (mNan) cc = hw_u2f(0x7fc00000u);
    return cc;
}

// UHistogramSimd::Fractile (UHistogramSimd.cpp:140-197)
__device__ float Fractile(const HistView& h, uint32_t lowOut, uint32_t total, float frac)
{
    const float nfz = frac * (float)total;
    uint32_t n = lowOut;
    if (n > 0 && (float)n >= nfz) return hw_u2f(0xff800000u);
    int i = 0;
    while ((n == 0 || (float)n < nfz) && i < NB) n = (n + h.Cnt(i++)) & 0xffffu;
    if ((float)n < nfz) return hw_u2f(0x7f800000u);
    i -= 1;
    const float x0 = h.Start(i);
    const uint32_t niz = h.Cnt(i);
    const int m = (int)n - (int)niz;
    return x0 + h.bs * (nfz - (float)m) / (float)(niz + 1);
}

__device__ __forceinline__ float Sum5Tree(const float* v) { return (v[0] + v[1]) + (v[2] + (v[3] + v[4])); }

enum { EM_PLAIN = 0, EM_LEARNED = 1 };

// Extra inputs/outputs of the EM kernel beyond the reference's.
struct EmExtra {
    const float* init;        // EM_LEARNED: network outputs [lane*64 + z][HWPA_MLP_OUT]
    const float* omap;        // EM_LEARNED: output map {off[6], scale[6], ...}
    uint8_t* fallback;        // EM_LEARNED: 1 = start rejected, model left for the plain EM
    float* diag;              // optional raw estimate [lane][HWPA_EM_DIAG_ROWS][64]
    const uint32_t* active;   // optional: number of valid ZMWs; lanes at or beyond it exit
    const uint8_t* laneSel;   // optional: only lanes whose 64 flags here are all set run
    float acceptConf;         // EM_LEARNED: minimum confidence to accept the capped EM
    float convTol;            // EM_LEARNED: relative log-likelihood convergence tolerance
    int32_t iterLimit;        // EM_LEARNED: iteration cap
    float gateLogit;          // EM_LEARNED: network confidence logit below which the capped EM is skipped
    int32_t laneMinConf;      // EM_LEARNED: a lane with fewer ZMWs above the gate gates all of them
};

// DmeEmHost::EstimateLaneDetModel (DmeEmHost.cpp:216-615).  EM_LEARNED replaces
// the start model of the EM loop (after PrelimEstimate and the scale factor,
// which still set the amplitude prior) by the network's prediction, caps the
// iterations, uses its own log-likelihood convergence tolerance and accepts
// the result only for converged ZMWs with status OK and enough confidence.
// exp(t) rounds to +0 for every t below this (exp(-103.98) is 2^-150).
constexpr float kExpZeroBelow = -105.0f;
// Boundaries per exact restart of the density recurrence of the E-step.
constexpr int kRecurBlock = 10;
// Bins between a count's load and its use in the E-step.
constexpr int kCntAhead = 2;

template <int MODE>
__global__ __launch_bounds__(64, 3) void DmeEstimateKernel(const hwpa_lane_hist* __restrict__ hists,
                                                        const hwpa_baseliner_stats* __restrict__ stats,
                                                        hwpa_lane_model* models, int32_t* iters, DmeDevCfg d,
                                                        float evolve, EmExtra x)
{
    const int lane = blockIdx.x;
    const int z = threadIdx.x;
    if (x.active && (uint32_t)lane * LS >= *x.active) return;
    if (x.laneSel && !__all(x.laneSel[(size_t)lane * LS + z] != 0)) return;
    bool gated = false;       // EM_LEARNED: the network's confidence sends this ZMW to the plain EM
    if (MODE == EM_LEARNED) {
        // Confidence gate: a ZMW below it takes no capped iteration of its own (it still
        // runs in step with its lane, but counts as done and is rejected); a lane with fewer
        // than laneMinConf ZMWs above it sends all of them to the plain EM (its iterations
        // would cost the whole lane for those few).  A lane gated whole leaves its models
        // as they are, as every rejected ZMW does, and stops here.
        gated = !(x.init[((size_t)lane * LS + z) * HWPA_MLP_OUT + 11] >= x.gateLogit);
        if (__popcll(__ballot(!gated)) < (uint32_t)x.laneMinConf) gated = true;
        if (__all(gated)) {
            x.fallback[(size_t)lane * LS + z] = 1;
            if (z == 0 && iters) iters[lane] = 0;
            return;
        }
    }
    // Tables of the E-step's log / reciprocal of bin probabilities, one entry per thread.
    __shared__ float sRc[HW_LOGTAB_N], sLc[HW_LOGTAB_N];
    static_assert(HW_LOGTAB_N == LS, "one table entry per thread");
    hw_logtab_entry(z, &sRc[z], &sLc[z]);
    __syncthreads();
    const hwpa_lane_hist& lh = hists[lane];
    const hwpa_baseliner_stats& bs = stats[lane];
    hwpa_lane_model& lm = models[lane];

    HistView h{&lh, z, lh.low_bound[z], lh.bin_size[z]};
    const uint32_t lowOut = lh.outlier_low[z], highOut = lh.outlier_high[z];
    int lastNz;
    const uint32_t inRange = InRange(h, &lastNz);
    const uint32_t numFrames = (inRange + ((lowOut + highOut) & 0xffffu)) & 0xffffu;

    const float rpa[NA] = {d.relAmp[0], d.relAmp[1], d.relAmp[2], d.relAmp[3]};
    const float rpaSqSum = (rpa[0] * rpa[0] + rpa[2] * rpa[2]) + (rpa[1] * rpa[1] + rpa[3] * rpa[3]);
    const float rpaMin = hw_minf(hw_minf(rpa[0], rpa[1]), hw_minf(rpa[2], rpa[3]));

    Model det = LoadModel(lm, z);
    det.conf *= evolve;

    // PrelimEstimate (DmeEmHost.cpp:161-214)
    {
        Model w = det;
        const float nBl = bs.baseline.m0[z];
        const float total = bs.full.basic.m0[z];
        const float blWeight = hw_maxf(nBl / total, 0.01f);
        const bool mask = nBl >= 3.0f;
        float blVar = mask ? Variance(bs.baseline.m0[z], bs.baseline.m1[z], bs.baseline.m2[z]) : w.bl.var;
        blVar = hw_clampf(blVar, d.blVarMin, d.blVarMax);
        const float blMean = mask ? (bs.baseline.m1[z] / bs.baseline.m0[z] + bs.baseline.offset[z]) : w.bl.mean;
        const float scale = hw_sqrtf(blVar / w.bl.var);
#pragma unroll
        for (int a = 0; a < NA; ++a) {
            const float mean = w.an[a].mean * scale;
            const float var = Covar(d, d.cv[a] * d.cv[a], mean, blVar);
            w.an[a].mean = mean;
            w.an[a].var = var;
            w.an[a].w = 0.25f * (1.0f - blWeight);
        }
        w.bl.w = blWeight;
        w.bl.mean = blMean;
        w.bl.var = blVar;
        float conf = 0.1f * hw_satlin(0.0f, 500.0f, nBl - 3.0f);
        conf *= hw_satlin(d.blVarMax, 0.5f * (d.blVarNominal + d.blVarMax), blVar);
        w.conf = conf;
        UpdateConf(d, det, w);
    }
    Model work = det;

    // PrelimScaleFactor (DmeEmHost.cpp:618-654) and ScaleModelSnr (:837-852)
    float scaleFactor;
    {
        const float bgSigma = hw_sqrtf(work.bl.var);
        const float thresh = 2.5f * bgSigma + work.bl.mean;
        const float bgCount = CumulativeCount(h, lowOut, inRange, thresh);
        const float totalCount = (float)((lowOut + inRange) & 0xffffu);
        const float fractile = 0.5f * (bgCount + totalCount) / totalCount;
        float avg = 0.0f;
#pragma unroll
        for (int a = 0; a < NA; ++a) avg += work.an[a].mean;
        avg /= (float)NA;
        const float w = hw_satlin(0.0f, d.scaleSnrConfTol, work.conf);
        float sf = Fractile(h, lowOut, numFrames, fractile) / avg;
        sf = (1.0f - w) * sf + w;
        scaleFactor = hw_minf(hw_maxf(sf, 0.1f), 10.0f);
    }
#pragma unroll
    for (int a = 0; a < NA; ++a) {
        work.an[a].mean = scaleFactor * work.an[a].mean;
        work.an[a].var = Covar(d, d.cv[a] * d.cv[a], work.an[a].mean, work.bl.var);
    }

    float rho[5], mu[5], var[5], rhoEst[5], muEst[5], varEst[5];
    const float rhoMin = 0.001f;
    rho[0] = hw_maxf(work.bl.w, rhoMin);
    mu[0] = work.bl.mean;
    var[0] = work.bl.var;
#pragma unroll
    for (int a = 0; a < NA; ++a) {
        rho[a + 1] = hw_maxf(work.an[a].w, rhoMin);
        mu[a + 1] = work.an[a].mean;
        var[a + 1] = work.an[a].var;
    }
    const float varQuant = (h.bs * h.bs) / 12.0f;
#pragma unroll
    for (int k = 0; k < 5; ++k) var[k] += varQuant;
    var[0] = hw_minf(var[0], d.blVarMax);
    {
        const float rs = Sum5Tree(rho);
#pragma unroll
        for (int k = 0; k < 5; ++k) rho[k] = rho[k] / rs;
    }
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        rhoEst[k] = rho[k];
        muEst[k] = mu[k];
        varEst[k] = var[k];
    }
    float s = ((mu[1] * rpa[0] + mu[2] * rpa[1]) + (mu[3] * rpa[2] + mu[4] * rpa[3])) / rpaSqSum;
    const float sExpect = s / scaleFactor;
    const float sExpectWeight = det.conf * d.pulseAmpReg;
    float logLike = -FLT_MAX;
    const float binProbMin = hw_maxf((float)numFrames * h.bs, 1.0f) * FLT_MIN;
    int status = (numFrames < 160u) ? HWPA_ZS_INSUF_DATA : HWPA_ZS_OK;
    bool converged = false;
    float mlLogLike = hw_u2f(0xff800000u);
    float njSum = 0.0f;
    for (int b = 0; b < NB; ++b) njSum += (float)h.Cnt(b);
    const float halfBin = 0.5f * h.bs;
    int iterLimit = d.iterLimit;
    if (MODE == EM_LEARNED) {
        // Network start: output j < 6 stands for off[j] + scale[j] * y[j] in units
        // of u = 2 binSize from the lower bound (means) and log(var / u^2);
        // outputs 6..10 are softmax mixture logits.
        const float* y = x.init + ((size_t)lane * LS + z) * HWPA_MLP_OUT;
        const float u = 2.0f * h.bs;
        float r[6];
#pragma unroll
        for (int j = 0; j < 6; ++j) r[j] = x.omap[j] + x.omap[6 + j] * y[j];
        float mx = y[6];
#pragma unroll
        for (int k = 1; k < 5; ++k) mx = hw_maxf(mx, y[6 + k]);
        float e[5], se = 0.0f;
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            e[k] = hw_expf(y[6 + k] - mx);
            se += e[k];
        }
#pragma unroll
        for (int k = 0; k < 5; ++k) rho[k] = hw_maxf(e[k] / se, rhoMin);
        const float rs = Sum5Tree(rho);
#pragma unroll
        for (int k = 0; k < 5; ++k) rho[k] = rho[k] / rs;
        mu[0] = h.lb + u * r[0];
        var[0] = hw_clampf((u * u) * hw_expf(r[1]), d.blVarMin, d.blVarMax);
#pragma unroll
        for (int a = 0; a < NA; ++a) {
            mu[a + 1] = h.lb + u * r[2 + a];
            var[a + 1] = Covar(d, d.cv[a] * d.cv[a], mu[a + 1], var[0]);
        }
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            rhoEst[k] = rho[k];
            muEst[k] = mu[k];
            varEst[k] = var[k];
        }
        s = ((mu[1] * rpa[0] + mu[2] * rpa[1]) + (mu[3] * rpa[2] + mu[4] * rpa[3])) / rpaSqSum;
        iterLimit = x.iterLimit;
    }

    int it = 0;
    for (; it < iterLimit; ++it) {
        // ---- E-step
        float varinv[5], pref[5], nhv[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            varinv[k] = 1.0f / var[k];
            pref[k] = (hw_cephes_logf(rho[k]) - 0.5f * hw_cephes_logf(var[k])) - 0.5f * d.log2pi;
            nhv[k] = -0.5f * varinv[k];
        }
        // Every use of a responsibility is boundaryProb * tau_i = exp(t_i)
        // (DmeEmHost.cpp:431-432,547-548), so the boundary carries the component
        // densities dd_i = exp(t_i) and boundaryProb is their sum.  The bin sums of
        // the reference, sum_b (dd_i(b) + dd_i(b+1)) n_b / binProb_b, are taken per
        // boundary: sum_c dd_i(c) (f_{c-1} + f_c) with f_b = n_b / binProb_b, one fused
        // multiply-add per sum.  The baseline second moment of the M-step is
        // accumulated in the same pass about the pivot piv (this iteration's baseline
        // mean): q1 and q2 are the first and second moments of (x - piv) under the
        // baseline weights, shifted to the new mean after the M-step.  Waves skip
        // exact zeros: exps of boundaries where every t_i < -105 (exp rounds to +0),
        // and the sums of boundaries whose two bins are empty for all 64 ZMWs.
        const float piv = mu[0];
        float probSum = 0.0f, wps = 0.0f, q1 = 0.0f, q2 = 0.0f;
        float ci[5] = {0.f, 0.f, 0.f, 0.f, 0.f}, mom1[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
        // Densities along the boundaries x_c = lb + c bs: t_i is quadratic in c, so
        // dd_i(c+1) = dd_i(c) R_i(c) with R_i(c) = exp(t_i(c+1) - t_i(c)) and
        // R_i(c+1) = R_i(c) Q_i, Q_i = exp(-bs^2 / var_i): two multiplies per density.
        // Every kRecurBlock boundaries dd_i and R_i are recomputed from exp.  The products
        // follow the ideal grid x_c + j bs; against exp at the rounded boundaries they
        // differ by |x - mu| / var times the boundary's rounding (below 1e-4 relative
        // within 14 sigma of the mean, 1e-5 near it).  A ZMW whose density at the
        // restart is below FLT_MIN while rising, or whose ratio exceeds e^80, takes exp
        // at every boundary of the block (a density coming out of underflow cannot be
        // continued from +0).
        float dd[5], rr[5], qq[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) qq[k] = hw_expf((2.0f * nhv[k]) * (h.bs * h.bs));
        uint32_t exMask = 0;
        bool anyEx = false;   // some ZMW of the wave takes exp at every boundary of this block
        auto tAt = [&](float x, int i) {
            const float y = x - mu[i];
            return hw_fmaf(y * y, nhv[i], pref[i]);
        };
        float nbs[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) nbs[k] = nhv[k] * h.bs;
        auto restart = [&](int c) {
            const float x0 = h.Start(c);
            exMask = 0;
#pragma unroll
            for (int i = 0; i < 5; ++i) {
                const float y0 = x0 - mu[i];
                const float t0 = hw_fmaf(y0 * y0, nhv[i], pref[i]);
                // t(x0 + bs) - t(x0) on the ideal grid (no cancellation of rounded t)
                const float dl = nbs[i] * hw_fmaf(2.0f, y0, h.bs);
                // exp(t) is +0 below the threshold: waves skip the exp when all ZMWs are, and
                // then the ratio's too.  The density stays +0 through the block (+0 times any
                // finite ratio >= 0); a rising one (dl > 0) takes exp at every boundary below.
                if (__all(t0 < kExpZeroBelow)) {
                    dd[i] = 0.0f;
                    rr[i] = 0.0f;
                } else {
                    dd[i] = hw_expf(t0);
                    rr[i] = hw_expf(dl);
                }
                const bool rec = (dd[i] >= FLT_MIN || dl <= 0.0f) && dl <= 80.0f;
                exMask |= (rec ? 0u : 1u) << i;
            }
            anyEx = __any(exMask != 0u);
        };
        auto advance = [&](int c) {
#pragma unroll
            for (int i = 0; i < 5; ++i) {
                dd[i] = dd[i] * rr[i];
                rr[i] = rr[i] * qq[i];
            }
            if (anyEx) {
                const float x = h.Start(c);
#pragma unroll
                for (int i = 0; i < 5; ++i)
                    if ((exMask >> i) & 1u) dd[i] = hw_expf(tAt(x, i));
            }
        };
        // Sums of boundary c with weight F = f_{c-1} + f_c; skipped when F is 0 for the
        // whole wave (every density is finite, so the terms are exact zeros).
        auto gather = [&](const float* dc, float x, float F) {
            if (__all(F == 0.0f)) return;
            const float G = x * F;
#pragma unroll
            for (int m = 0; m < 5; ++m) {
                ci[m] = hw_fmaf(dc[m], F, ci[m]);
                mom1[m] = hw_fmaf(dc[m], G, mom1[m]);
            }
            const float dev = x - piv;
            const float w = dc[0] * F;
            q1 = hw_fmaf(w, dev, q1);
            q2 = hw_fmaf(w * dev, dev, q2);
        };
        restart(0);
        float bpPrev = Sum5Tree(dd), dPrev[5];
#pragma unroll
        for (int m = 0; m < 5; ++m) dPrev[m] = dd[m];
        float fPrev = 0.0f;
        // Bin counts kCntAhead bins ahead, so their loads are in flight over that many bins.
        uint32_t cq[kCntAhead];
#pragma unroll
        for (int j = 0; j < kCntAhead; ++j) cq[j] = h.Cnt(j);
        static_assert(NB % kRecurBlock == 0, "whole blocks of boundaries");
        // The tail of the histogram is left out once, for every ZMW of the wave, the bins
        // ahead are empty, the boundary lies beyond every mode's mean (so each density
        // only falls from here on) and bs times the density sum is below 2^-27 of
        // probSum, as is binProbMin: then every remaining bin probability is below half
        // an ulp of probSum, so adding them would leave it unchanged, and every other sum
        // of those boundaries is an exact zero (empty bins).  The result equals the full
        // pass bit for bit.
        float maxMu = mu[0];
#pragma unroll
        for (int k = 1; k < 5; ++k) maxMu = hw_maxf(maxMu, mu[k]);
        int bEnd = NB;
        for (int b0 = 0; b0 < NB; b0 += kRecurBlock) {
            if (b0 > 0) {
                const float tol = hw_ldexpf(probSum, -27);
                const bool tail = lastNz < b0 && h.Start(b0) > maxMu && h.bs * bpPrev < tol && binProbMin < tol;
                if (__all(tail)) {
                    bEnd = b0;
                    break;
                }
            }
#pragma unroll
            for (int j = 0; j < kRecurBlock; ++j) {
                const int b = b0 + j;
                const uint32_t cnt = cq[0];
#pragma unroll
                for (int q = 0; q + 1 < kCntAhead; ++q) cq[q] = cq[q + 1];
                cq[kCntAhead - 1] = h.Cnt(b + kCntAhead);
                if (j == kRecurBlock - 1) restart(b + 1);
                else advance(b + 1);
                const float bpNext = Sum5Tree(dd);
                // binp >= binProbMin >= FLT_MIN and finite: the log needs no special cases
                const float binp = hw_max_pos(halfBin * (bpPrev + bpNext), binProbMin);
                probSum += binp;
                float f = 0.0f;   // n_b / binp_b, +0 for an empty bin
                if (!__all(cnt == 0u)) {
                    const float nj = (float)cnt;
                    float lg, rcp;
                    hw_log_rcp_tab(binp, sRc, sLc, &lg, &rcp);
                    wps = hw_fmaf(nj, lg, wps);
                    f = nj * rcp;
                }
                gather(dPrev, h.Start(b), fPrev + f);
                bpPrev = bpNext;
#pragma unroll
                for (int m = 0; m < 5; ++m) dPrev[m] = dd[m];
                fPrev = f;
            }
        }
        gather(dPrev, h.Start(bEnd), fPrev);
        const float q0 = ci[0];
#pragma unroll
        for (int m = 0; m < 5; ++m) {
            ci[m] *= halfBin;
            mom1[m] *= halfBin;
        }
        const float logLikePrev = logLike;
        logLike = wps - njSum * hw_cephes_logf(probSum);
        const float ds = s - sExpect;
        logLike -= 0.5f * (ds * ds) * sExpectWeight;
        const float delta = logLike - logLikePrev;
        const float convTol = (MODE == EM_LEARNED ? x.convTol : 1e-4f) * hw_fabsf(logLike);
        bool conv = (delta >= 0.0f) && (delta < convTol);
        conv = conv && (logLike >= mlLogLike);
        if (conv) {
            mlLogLike = logLike;
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                rhoEst[k] = rho[k];
                muEst[k] = mu[k];
                varEst[k] = var[k];
            }
        }
        converged = converged || conv;
        // Lane-wide early exit (DmeEmHost.cpp:485)
        const bool done = converged || (status != HWPA_ZS_OK) || gated;
        if (!d.iterToLimit && __all(done)) break;

        // ---- M-step
        float c[5];
#pragma unroll
        for (int m = 0; m < 5; ++m) c[m] = hw_maxf(ci[m], 0.1f);
#pragma unroll
        for (int m = 0; m < 5; ++m) rho[m] = c[m] / njSum;
        {
            const float rs = Sum5Tree(rho);
#pragma unroll
            for (int m = 0; m < 5; ++m) rho[m] = rho[m] / rs;
        }
        mu[0] = mom1[0] / c[0];
        float numer = 0.0f, denom = 0.0f;
#pragma unroll
        for (int a = 0; a < NA; ++a) {
            const int i = a + 1;
            numer += rpa[a] * varinv[i] * mom1[i];
            denom += (rpa[a] * rpa[a]) * varinv[i] * c[i];
        }
        numer += sExpect * sExpectWeight;
        denom += sExpectWeight;
        s = numer / denom;
        const float minPulseMean = hw_maxf(1.0f, d.signalScaler);
        const float sMin = hw_maxf(0.5f * hw_sqrtf(var[0]), minPulseMean) / rpaMin;
        if (s < sMin) {
            s = sMin;
            status |= HWPA_ZS_VLOW_SIGNAL;
        }
#pragma unroll
        for (int a = 0; a < NA; ++a) mu[a + 1] = s * rpa[a];
        // Baseline second moment about the new mean, from the moments about piv.
        const float dl = mu[0] - piv;
        float mom2 = (q2 - (2.0f * dl) * q1) + (dl * dl) * q0;
        mom2 *= 0.5f * h.bs;
        var[0] = mom2 / c[0] + varQuant;
        var[0] = hw_clampf(var[0], d.blVarMin, d.blVarMax);
#pragma unroll
        for (int a = 0; a < NA; ++a) var[a + 1] = Covar(d, d.cv[a] * d.cv[a], mu[a + 1], var[0]);
    }
    if (!converged) status |= HWPA_ZS_NO_CONVERGE;
    if (z == 0 && iters) iters[lane] = it;

    work.bl.w = rhoEst[0];
    work.bl.mean = muEst[0];
    work.bl.var = varEst[0];
#pragma unroll
    for (int a = 0; a < NA; ++a) {
        work.an[a].w = rhoEst[a + 1];
        work.an[a].mean = muEst[a + 1];
        work.an[a].var = varEst[a + 1];
    }

    // ComputeConfidence (DmeEmHost.cpp:717-803); the G-test factor stays 1.
    float conf = 1.0f;
    {
        float cf[8];
        cf[0] = converged ? 1.0f : 0.0f;
        cf[1] = hw_satlin(0.05f, 0.15f, work.bl.w);
        float x = (work.bl.mean * work.bl.mean) / work.bl.var;
        cf[2] = hw_cephes_expf(-x / (2.0f * (1.0f * 1.0f)));
        x = hw_dlog2f(work.bl.var / det.bl.var);
        const float tol = 1.5f / (0.5f + det.conf);
        cf[3] = hw_cephes_expf(-x * x / (2.0f * (tol * tol)));
        x = 1.0f;
        if (d.mixThresh1 > 0.0f) {
            x *= hw_satlin(d.mixThresh0, d.mixThresh1, work.an[0].w);
            x *= hw_satlin(d.mixThresh0, d.mixThresh1, work.an[NA - 1].w);
        }
        cf[4] = x;
        const float bgSigma = hw_sqrtf(work.bl.var);
        x = work.an[NA - 1].mean;
        x /= bgSigma;
        cf[5] = hw_satlin(d.snrThresh0, d.snrThresh1, x);
        if (d.snrDrop < 0.0f) {
            cf[6] = 1.0f;
        } else {
            const float snrEst = work.an[0].mean / bgSigma;
            const float refBgSigma = hw_sqrtf(det.bl.var);
            const float refSnr0 = det.an[0].mean / refBgSigma;
            float refSnr1 = det.an[1].mean / refBgSigma;
            refSnr1 *= d.snrDrop;
            refSnr1 *= hw_minf(det.conf, 1.0f);
            cf[6] = hw_satlin(refSnr1, hw_sqrtf(refSnr0 * refSnr1), snrEst);
        }
        cf[7] = 1.0f;
#pragma unroll
        for (int k = 0; k < 8; ++k) conf *= cf[k];
    }
    conf = hw_minf(hw_maxf(0.0f, conf), 1.0f);
    conf = (conf >= d.successThresh) ? conf : 0.0f;
    if (x.diag) {
        float* dg = x.diag + (size_t)lane * HWPA_EM_DIAG_ROWS * LS + z;
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            dg[k * LS] = rhoEst[k];
            dg[(5 + k) * LS] = muEst[k];
            dg[(10 + k) * LS] = varEst[k];
        }
        dg[15 * LS] = conf;
        dg[16 * LS] = converged ? 1.0f : 0.0f;
        dg[17 * LS] = (float)status;
    }
    if (MODE == EM_LEARNED) {
        const bool accept = !gated && converged && status == HWPA_ZS_OK && conf >= x.acceptConf;
        x.fallback[(size_t)lane * LS + z] = accept ? 0 : 1;
        if (!accept) return;
    }
    work.conf = conf;
    UpdateConf(d, det, work);
    StoreModel(det, lm, z);
}

// ---- fallback compaction for the learned start ---------------------------------
// A lane whose 64 ZMWs were all rejected runs the plain EM in place.  The
// rejected ZMWs of the other lanes are packed densely (in lane order, so the
// grouping is deterministic) into compact lanes, run through the plain EM and
// scattered back.

__device__ __forceinline__ void ColStatAcc(const hwpa_stat_acc& s, hwpa_stat_acc& d, int sz, int dz)
{
    d.offset[dz] = s.offset[sz];
    d.m0[dz] = s.m0[sz];
    d.m1[dz] = s.m1[sz];
    d.m2[dz] = s.m2[sz];
}

__device__ void ColStats(const hwpa_baseliner_stats& s, hwpa_baseliner_stats& d, int sz, int dz)
{
    ColStatAcc(s.full.basic, d.full.basic, sz, dz);
    d.full.lag_m2[dz] = s.full.lag_m2[sz];
    for (int k = 0; k < HWPA_AUTOCORR_LAG; ++k) {
        d.full.fbuf[k][dz] = s.full.fbuf[k][sz];
        d.full.bbuf[k][dz] = s.full.bbuf[k][sz];
    }
    d.full.bidx[0][dz] = s.full.bidx[0][sz];
    d.full.bidx[1][dz] = s.full.bidx[1][sz];
    d.trace_min[dz] = s.trace_min[sz];
    d.trace_max[dz] = s.trace_max[sz];
    ColStatAcc(s.baseline, d.baseline, sz, dz);
    d.raw_baseline_sum[dz] = s.raw_baseline_sum[sz];
    ColStatAcc(s.background, d.background, sz, dz);
}

__device__ void ColHist(const hwpa_lane_hist& s, hwpa_lane_hist& d, int sz, int dz)
{
    d.low_bound[dz] = s.low_bound[sz];
    d.bin_size[dz] = s.bin_size[sz];
    d.outlier_low[dz] = s.outlier_low[sz];
    d.outlier_high[dz] = s.outlier_high[sz];
    for (int b = 0; b < NB; ++b) d.bin_count[b][dz] = s.bin_count[b][sz];
}

__device__ void ColModel(const hwpa_lane_model& s, hwpa_lane_model& d, int sz, int dz)
{
    for (int a = 0; a < NA; ++a) {
        d.analogs[a].means[dz] = s.analogs[a].means[sz];
        d.analogs[a].vars[dz] = s.analogs[a].vars[sz];
        d.analogs[a].weights[dz] = s.analogs[a].weights[sz];
    }
    d.baseline.means[dz] = s.baseline.means[sz];
    d.baseline.vars[dz] = s.baseline.vars[sz];
    d.baseline.weights[dz] = s.baseline.weights[sz];
    d.confidence[dz] = s.confidence[sz];
}

__global__ __launch_bounds__(64) void FallbackCountKernel(const uint8_t* __restrict__ fb, uint32_t* __restrict__ cnt)
{
    const uint64_t m = __ballot(fb[(size_t)blockIdx.x * LS + threadIdx.x] != 0);
    if (threadIdx.x == 0) cnt[blockIdx.x] = (uint32_t)__popcll(m);
}

// Exclusive scan of per-lane counts in one workgroup (fixed order).  Lanes whose
// 64 ZMWs all fall back run the plain EM in place and take no compact slots:
// total[0] counts the compacted ZMWs, total[1] every fallback ZMW.
__global__ __launch_bounds__(1024) void FallbackScanKernel(const uint32_t* __restrict__ cnt, uint32_t lanes,
                                                           uint32_t* __restrict__ off, uint32_t* __restrict__ total)
{
    __shared__ uint32_t part[1024], whole[1024];
    const uint32_t per = (lanes + 1023) / 1024;
    const uint32_t b0 = threadIdx.x * per, b1 = min(lanes, b0 + per);
    uint32_t s = 0, w = 0;
    for (uint32_t i = b0; i < b1; ++i) {
        const uint32_t c = cnt[i];
        s += c == LS ? 0u : c;
        w += c == LS ? c : 0u;
    }
    part[threadIdx.x] = s;
    whole[threadIdx.x] = w;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0, all = 0;
        for (int i = 0; i < 1024; ++i) {
            const uint32_t v = part[i];
            part[i] = run;
            run += v;
            all += whole[i];
        }
        total[0] = run;
        total[1] = run + all;
    }
    __syncthreads();
    uint32_t run = part[threadIdx.x];
    for (uint32_t i = b0; i < b1; ++i) {
        off[i] = run;
        const uint32_t c = cnt[i];
        run += c == LS ? 0u : c;
    }
}

__global__ __launch_bounds__(64) void FallbackGatherKernel(const uint8_t* __restrict__ fb, const uint32_t* __restrict__ off,
                                                           const hwpa_lane_hist* __restrict__ hist,
                                                           const hwpa_baseliner_stats* __restrict__ stats,
                                                           const hwpa_lane_model* __restrict__ models,
                                                           hwpa_lane_hist* cHist, hwpa_baseliner_stats* cStats,
                                                           hwpa_lane_model* cModels, uint32_t* __restrict__ map)
{
    const int lane = blockIdx.x, z = threadIdx.x;
    const bool f = fb[(size_t)lane * LS + z] != 0;
    const uint64_t m = __ballot(f);
    if (!f || m == ~0ull) return;
    const uint32_t rank = (uint32_t)__popcll(m & ((1ull << z) - 1ull));
    const uint32_t pos = off[lane] + rank;
    const uint32_t cl = pos / LS;
    const int dz = (int)(pos % LS);
    ColHist(hist[lane], cHist[cl], z, dz);
    ColStats(stats[lane], cStats[cl], z, dz);
    ColModel(models[lane], cModels[cl], z, dz);
    map[pos] = (uint32_t)lane * LS + (uint32_t)z;
}

// Fills the unused slots of the last, partly filled compact lane with copies of
// its first ZMW so that its EM behaves like a full lane.
__global__ __launch_bounds__(64) void FallbackPadKernel(const uint32_t* __restrict__ total, hwpa_lane_hist* cHist,
                                                        hwpa_baseliner_stats* cStats, hwpa_lane_model* cModels)
{
    const uint32_t t = *total;
    const uint32_t used = t % LS;
    if (used == 0) return;
    const uint32_t cl = t / LS;
    const int z = threadIdx.x;
    if ((uint32_t)z < used) return;
    ColHist(cHist[cl], cHist[cl], 0, z);
    ColStats(cStats[cl], cStats[cl], 0, z);
    ColModel(cModels[cl], cModels[cl], 0, z);
}

__global__ __launch_bounds__(64) void FallbackScatterKernel(const uint32_t* __restrict__ total,
                                                            const uint32_t* __restrict__ map,
                                                            const hwpa_lane_model* __restrict__ cModels,
                                                            hwpa_lane_model* models)
{
    const uint32_t pos = blockIdx.x * LS + threadIdx.x;
    if (pos >= *total) return;
    const uint32_t dst = map[pos];
    ColModel(cModels[blockIdx.x], models[dst / LS], (int)threadIdx.x, (int)(dst % LS));
}

int32_t DmeInitModels(const hwpa_config& cfg, const hwpa_baseliner_stats* stats, hwpa_lane_model* models,
                      uint32_t lanes, hipStream_t s)
{
    if (!stats || !models || lanes == 0) return Fail(HWPA_ERR_INVALID_ARG, "dme_init_models: bad argument");
    hipLaunchKernelGGL(DmeInitKernel, dim3(lanes), dim3(LS), 0, s, stats, models, MakeDevCfg(cfg));
    return CheckLaunch("DmeInitKernel");
}

// DmeEmHost::EvolveModel (DmeEmHost.cpp:806-834), DetectionModelHost::EvolveConfidence
static float EvolveFactor(const DmeDevCfg& d, const int32_t model_fi[2], const int32_t est_fi[2])
{
    const int32_t tModel = model_fi[0] + (int32_t)((uint32_t)(model_fi[1] - model_fi[0]) / 2u);
    const int32_t tEst = est_fi[0] + (int32_t)((uint32_t)(est_fi[1] - est_fi[0]) / 2u);
    const float thl = 0.5f * (float)(tModel + tEst);
    const float hl = d.hl0 + (d.hl1 - d.hl0) * hw_satlin(56160.0f, 168480.0f, thl);
    const float tDiff = (float)(tEst - tModel);
    return hw_dexp2f(-hw_fabsf(tDiff) / hl);
}

int32_t DmeEstimate(const hwpa_config& cfg, const hwpa_lane_hist* hist, const hwpa_baseliner_stats* stats,
                    hwpa_lane_model* models, const int32_t model_fi[2], const int32_t est_fi[2], uint32_t lanes,
                    int32_t* d_iters, hipStream_t s, float* diag)
{
    if (!hist || !stats || !models || lanes == 0) return Fail(HWPA_ERR_INVALID_ARG, "dme_estimate: bad argument");
    const DmeDevCfg d = MakeDevCfg(cfg);
    const float evolve = EvolveFactor(d, model_fi, est_fi);
    EmExtra x{};
    x.diag = diag;
    hipLaunchKernelGGL(DmeEstimateKernel<EM_PLAIN>, dim3(lanes), dim3(LS), 0, s, hist, stats, models, d_iters, d,
                       evolve, x);
    return CheckLaunch("DmeEstimateKernel");
}

void LearnedWsFree(LearnedWs& w)
{
    DeviceFree(w.y);
    DeviceFree(w.fallback);
    DeviceFree(w.cnt);
    DeviceFree(w.off);
    DeviceFree(w.total);
    DeviceFree(w.map);
    DeviceFree(w.cHist);
    DeviceFree(w.cStats);
    DeviceFree(w.cModels);
    DeviceFree(w.cIters);
    w = LearnedWs{};
}

int32_t LearnedWsAlloc(LearnedWs& w, uint32_t lanes)
{
    LearnedWsFree(w);
    const size_t Z = (size_t)lanes * LS;
    struct {
        void** p;
        size_t n;
    } req[] = {
        {(void**)&w.y, Z * HWPA_MLP_OUT * 4},
        {(void**)&w.fallback, Z},                 {(void**)&w.cnt, (size_t)lanes * 4},
        {(void**)&w.off, (size_t)lanes * 4},      {(void**)&w.total, 16},
        {(void**)&w.map, Z * 4},                  {(void**)&w.cHist, (size_t)lanes * sizeof(hwpa_lane_hist)},
        {(void**)&w.cStats, (size_t)lanes * sizeof(hwpa_baseliner_stats)},
        {(void**)&w.cModels, (size_t)lanes * sizeof(hwpa_lane_model)}, {(void**)&w.cIters, (size_t)lanes * 4},
    };
    for (auto& r : req) {
        const int32_t rc = DeviceAlloc(r.p, r.n, "learned-start workspace");
        if (rc != HWPA_OK) {
            LearnedWsFree(w);
            return rc;
        }
    }
    w.lanes = lanes;
    return HWPA_OK;
}

int32_t DmeEstimateLearned(const hwpa_config& cfg, hwpa_initializer* mlp, LearnedWs& w, const hwpa_lane_hist* hist,
                           const hwpa_baseliner_stats* stats, hwpa_lane_model* models, const int32_t model_fi[2],
                           const int32_t est_fi[2], uint32_t lanes, int32_t* d_iters, hipStream_t s)
{
    if (!mlp || !hist || !stats || !models || lanes == 0 || lanes > w.lanes)
        return Fail(HWPA_ERR_INVALID_ARG, "dme_estimate_learned: bad argument");
    if (cfg.learned_em_iteration_limit == 0) return Fail(HWPA_ERR_BAD_CONFIG, "learned EM iteration limit must be positive");
    const DmeDevCfg d = MakeDevCfg(cfg);
    const float evolve = EvolveFactor(d, model_fi, est_fi);
    int32_t rc = MlpForwardFromHist(mlp, hist, lanes, w.y, s);
    if (rc != HWPA_OK) return rc;
    EmExtra x{};
    x.init = w.y;
    x.omap = MlpOutputMap(mlp);
    x.fallback = w.fallback;
    x.acceptConf = cfg.learned_conf_threshold;
    x.convTol = cfg.learned_conv_tol;
    x.iterLimit = (int32_t)cfg.learned_em_iteration_limit;
    x.gateLogit = MlpGateLogit(cfg.learned_gate_threshold);
    x.laneMinConf = (int32_t)cfg.learned_lane_min_confident;
    hipLaunchKernelGGL(DmeEstimateKernel<EM_LEARNED>, dim3(lanes), dim3(LS), 0, s, hist, stats, models, d_iters, d,
                       evolve, x);
    if ((rc = CheckLaunch("DmeEstimateKernel")) != HWPA_OK) return rc;
    hipLaunchKernelGGL(FallbackCountKernel, dim3(lanes), dim3(LS), 0, s, w.fallback, w.cnt);
    if ((rc = CheckLaunch("FallbackCountKernel")) != HWPA_OK) return rc;
    hipLaunchKernelGGL(FallbackScanKernel, dim3(1), dim3(1024), 0, s, w.cnt, lanes, w.off, w.total);
    if ((rc = CheckLaunch("FallbackScanKernel")) != HWPA_OK) return rc;
    hipLaunchKernelGGL(FallbackGatherKernel, dim3(lanes), dim3(LS), 0, s, w.fallback, w.off, hist, stats, models,
                       w.cHist, w.cStats, w.cModels, w.map);
    if ((rc = CheckLaunch("FallbackGatherKernel")) != HWPA_OK) return rc;
    hipLaunchKernelGGL(FallbackPadKernel, dim3(1), dim3(LS), 0, s, w.total, w.cHist, w.cStats, w.cModels);
    if ((rc = CheckLaunch("FallbackPadKernel")) != HWPA_OK) return rc;
    EmExtra q{};
    q.laneSel = w.fallback;
    hipLaunchKernelGGL(DmeEstimateKernel<EM_PLAIN>, dim3(lanes), dim3(LS), 0, s, hist, stats, models,
                       static_cast<int32_t*>(nullptr), d, evolve, q);
    if ((rc = CheckLaunch("DmeEstimateKernel")) != HWPA_OK) return rc;
    EmExtra p{};
    p.active = w.total;
    hipLaunchKernelGGL(DmeEstimateKernel<EM_PLAIN>, dim3(lanes), dim3(LS), 0, s, w.cHist, w.cStats, w.cModels, w.cIters,
                       d, evolve, p);
    if ((rc = CheckLaunch("DmeEstimateKernel")) != HWPA_OK) return rc;
    hipLaunchKernelGGL(FallbackScatterKernel, dim3(lanes), dim3(LS), 0, s, w.total, w.map, w.cModels, models);
    return CheckLaunch("FallbackScatterKernel");
}

}  // namespace hwpa

// ---- learned-start estimation without allocation ---------------------------
// The workspace of DmeEstimateLearned for at most HWPA_NOALLOC_MAX_LANES lanes,
// in one device array.  Host code never passes `scratch` itself to a kernel
// (its host-side name is not a device address): NoAllocScratch looks the
// array's device address up once per device with hipGetSymbolAddress.
// Every buffer carved from it is written before it is read within one call,
// so nothing needs clearing between calls.

namespace hwpa {

constexpr size_t kNoAllocAlign = 256;
constexpr size_t AlignUp(size_t n) { return (n + kNoAllocAlign - 1) / kNoAllocAlign * kNoAllocAlign; }

// Bytes of the learned-start workspace for `lanes` lanes, each buffer
// starting on a 256-byte boundary (the order of CarveLearnedWs).
constexpr size_t LearnedWsBytes(size_t lanes)
{
    return AlignUp(lanes * LS * HWPA_MLP_OUT * 4) + AlignUp(lanes * LS) + AlignUp(lanes * 4) + AlignUp(lanes * 4) +
           AlignUp(16) + AlignUp(lanes * LS * 4) + AlignUp(lanes * sizeof(hwpa_lane_hist)) +
           AlignUp(lanes * sizeof(hwpa_baseliner_stats)) + AlignUp(lanes * sizeof(hwpa_lane_model)) +
           AlignUp(lanes * 4);
}

constexpr size_t kNoAllocBytes = LearnedWsBytes(HWPA_NOALLOC_MAX_LANES);

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kNoAllocMaxDevices = 64;

static int32_t NoAllocScratch(unsigned char** out)
{
    static std::mutex mu;
    static void* addr[kNoAllocMaxDevices] = {};
    int dev = -1;
    const hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return Fail(HWPA_ERR_HIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= kNoAllocMaxDevices) return Fail(HWPA_ERR_UNSUPPORTED, "device index beyond the workspace table");
    std::lock_guard<std::mutex> g(mu);
    if (!addr[dev]) {
        void* p = nullptr;
        const hipError_t r = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (r != hipSuccess || !p)
            return Fail(HWPA_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(r));
        addr[dev] = p;
    }
    *out = static_cast<unsigned char*>(addr[dev]);
    return HWPA_OK;
}

static LearnedWs CarveLearnedWs(unsigned char* base, uint32_t lanes)
{
    LearnedWs w{};
    size_t at = 0;
    auto take = [&](size_t n) {
        unsigned char* p = base + at;
        at += AlignUp(n);
        return p;
    };
    const size_t L = lanes;
    w.y = reinterpret_cast<float*>(take(L * LS * HWPA_MLP_OUT * 4));
    w.fallback = take(L * LS);
    w.cnt = reinterpret_cast<uint32_t*>(take(L * 4));
    w.off = reinterpret_cast<uint32_t*>(take(L * 4));
    w.total = reinterpret_cast<uint32_t*>(take(16));
    w.map = reinterpret_cast<uint32_t*>(take(L * LS * 4));
    w.cHist = reinterpret_cast<hwpa_lane_hist*>(take(L * sizeof(hwpa_lane_hist)));
    w.cStats = reinterpret_cast<hwpa_baseliner_stats*>(take(L * sizeof(hwpa_baseliner_stats)));
    w.cModels = reinterpret_cast<hwpa_lane_model*>(take(L * sizeof(hwpa_lane_model)));
    w.cIters = reinterpret_cast<int32_t*>(take(L * 4));
    w.lanes = lanes;
    return w;
}

}  // namespace hwpa

extern "C" int32_t hwpa_dme_estimate_learned_noalloc(const hwpa_config* cfg, hwpa_initializer* m,
                                                     const hwpa_lane_hist* d_hist,
                                                     const hwpa_baseliner_stats* d_stats, hwpa_lane_model* d_models,
                                                     const int32_t model_fi[2], const int32_t est_fi[2],
                                                     uint32_t lanes, int32_t* d_iters, uint32_t* fallback_count,
                                                     void* stream)
{
    using namespace hwpa;
    if (!cfg || !m || !model_fi || !est_fi || lanes == 0)
        return Fail(HWPA_ERR_INVALID_ARG, "dme_estimate_learned_noalloc: null argument or no lanes");
    if (lanes > HWPA_NOALLOC_MAX_LANES)
        return Fail(HWPA_ERR_INVALID_ARG, "dme_estimate_learned_noalloc: " + std::to_string(lanes) +
                                              " lanes exceed the fixed workspace of " +
                                              std::to_string(HWPA_NOALLOC_MAX_LANES) + " lanes");
    unsigned char* base = nullptr;
    int32_t rc = NoAllocScratch(&base);
    if (rc != HWPA_OK) return rc;
    LearnedWs w = CarveLearnedWs(base, lanes);
    hipStream_t s = AsStream(stream);
    rc = DmeEstimateLearned(*cfg, m, w, d_hist, d_stats, d_models, model_fi, est_fi, lanes, d_iters, s);
    if (hipStreamSynchronize(s) != hipSuccess && rc == HWPA_OK) rc = Fail(HWPA_ERR_HIP, "learned EM failed");
    if (rc == HWPA_OK && fallback_count) {
        if (hipMemcpy(fallback_count, w.total + 1, 4, hipMemcpyDeviceToHost) != hipSuccess)
            rc = Fail(HWPA_ERR_HIP, "fallback count copy failed");
    }
    return rc;
}
