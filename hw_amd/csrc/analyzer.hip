// analyzer.hip -- the per-pool analyzer behind the C-ABI boundary
// (DynamicEstimateBatchAnalyzer, analyzer/BatchAnalyzer.cpp:100-118,149-170,288-373;
// DetectionModelEstimator::AddBatch, traceAnalysis/DetectionModelEstimator.cpp:93-160).
//
// The host side only sequences stages and tracks frame intervals; every
// stage runs on the GPU on the analyzer's own stream, with all per-pool
// state resident in HBM between chunks.  Chunk order and pool identity are
// checked before any work is enqueued.

#include "internal.h"

#include <algorithm>
#include <cstring>
#include <vector>

namespace hwpa {

__global__ void ZeroCountsKernel(uint32_t* counts, uint32_t* ovf, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        counts[i] = 0;
        ovf[i] = 0;
    }
}

struct Interval {
    int32_t lo = 0, hi = 0;
    bool Empty() const { return lo == hi; }
    void Hull(int32_t a, int32_t b)
    {
        if (Empty()) {
            lo = a;
            hi = b;
        } else {
            lo = std::min(lo, a);
            hi = std::max(hi, b);
        }
    }
};

enum { kHistInit = 0, kDmeInit = 1, kSequencing = 2 };
enum { kStBaseline = 0, kStAggregate, kStHistogram, kStDme, kStLabels, kStPulses, kStMetrics, kNumStages };
// Kernel groups one process call enqueued; equal plans enqueue the same launch sequence.
enum : uint32_t {
    kPlanCopyIn = 1u << 0, kPlanEarlyZero = 1u << 1, kPlanAggregate = 1u << 2, kPlanHistAdd = 1u << 3,
    kPlanDmeInit = 1u << 4, kPlanEstimate = 1u << 5, kPlanLearned = 1u << 6, kPlanCapture = 1u << 7,
    kPlanReset = 1u << 8, kPlanLabels = 1u << 9, kPlanPulses = 1u << 10, kPlanLateZero = 1u << 11,
    kPlanEvents = 1u << 12, kPlanMetrics = 1u << 13, kPlanUint8 = 1u << 15,
    kPlanEmptyPulses = 1u << 16
};

// FixedModelBatchAnalyzer ctor (analyzer/BatchAnalyzer.cpp:120-145) with
// StaticDetModelConfig::SetupAnalogs (dataTypes/configs/StaticDetModelConfig.h:53-68):
// refSignal = refSnr sqrt(blVar); mean_a = blMean + relAmp_a refSignal;
// var_a = blVar + mean_a + (excessNoiseCV_a mean_a)^2 (the square as one product).
static void StaticModel(const hwpa_config& c, hwpa_lane_model* m)
{
    std::memset(m, 0, sizeof(*m));
    const float blMean = c.static_baseline_mean, blVar = c.static_baseline_variance;
    const float refSignal = c.movie.ref_snr * std::sqrt(blVar);
    for (int z = 0; z < LS; ++z) {
        m->baseline.means[z] = hw_f2h(blMean);
        m->baseline.vars[z] = hw_f2h(blVar);
    }
    for (int a = 0; a < NA; ++a) {
        const float mean = blMean + c.movie.analogs[a].rel_amplitude * refSignal;
        const float cvm = c.movie.analogs[a].excess_noise_cv * mean;
        const float var = blVar + mean + cvm * cvm;
        for (int z = 0; z < LS; ++z) {
            m->analogs[a].means[z] = hw_f2h(mean);
            m->analogs[a].vars[z] = hw_f2h(var);
        }
    }
}

}  // namespace hwpa

using namespace hwpa;

struct hwpa_analyzer {
    hwpa_config cfg;
    uint32_t poolId, lanes, frames;
    int32_t nextFrame = 0;
    uint32_t baselinerLatency = 0, dmeDelay = 0;
    bool fullEstimation = false;   // DynamicEstimate: a full estimation has occurred
    bool modelInitialized = false; // SingleEstimate: the first estimate is in place
    int status = kHistInit;
    int32_t framesRemaining = 0;
    Interval aggFi, histFi, modelFi;
    int32_t flags[4] = {0, 0, 0, 0};
    hipStream_t stream = nullptr;
    // Host-buffer input: chunks are copied on their own stream into two staging buffers in
    // turn, so the copy of one chunk runs while the previous chunk is processed.
    hipStream_t copyStream = nullptr;
    int16_t* dTraces2 = nullptr;
    hipEvent_t copied[2] = {}, bufFree[2] = {};
    bool bufUsed[2] = {};
    int copyBuf = 0;
    int pendingCopy = -1;   // staging buffer whose copy the current call waits for before returning
    bool profiling = false;
    hipEvent_t ev[kNumStages + 1] = {};
    bool stageRan[kNumStages] = {};

    BaselinerDev* bl = nullptr;
    LabelerDev* lab = nullptr;
    PulserDev* pul = nullptr;
    hwpa_metrics* met = nullptr;
    hwpa_basecalling_metrics* dMetrics = nullptr;
    int16_t* dTraces = nullptr;
    int16_t* dBaselined = nullptr;
    hwpa_baseliner_stats* dStats = nullptr;
    hwpa_baseliner_stats* dAgg = nullptr;
    hwpa_lane_hist* dHist = nullptr;
    void* dEdge = nullptr;
    hwpa_lane_model* dModels = nullptr;
    int16_t* dLabels = nullptr;
    int16_t* dLatent = nullptr;
    float* dScore = nullptr;
    hwpa_pulse* dPulses = nullptr;
    uint32_t* dCounts = nullptr;
    uint32_t* dOvf = nullptr;
    hwpa_stat_acc* dPulseBl = nullptr;
    int32_t* dIters = nullptr;
    bool ranEstimate = false;
    hwpa_initializer* mlp = nullptr;
    LearnedWs lws;
    bool ranLearned = false;
    uint32_t plan = 0;  // work enqueued by the last process call (kPlan* bits)
    // training-pair capture
    uint16_t* capX = nullptr;
    float* capT = nullptr;
    uint8_t* capValid = nullptr;
    float capMinConf = 0.0f;
    float* dDiag = nullptr;
};

static void FreeAnalyzer(hwpa_analyzer* a)
{
    if (!a) return;
    if (a->copyStream) (void)hipStreamSynchronize(a->copyStream);
    if (a->stream) (void)hipStreamSynchronize(a->stream);
    BaselinerDestroy(a->bl);
    LabelerDestroy(a->lab);
    PulserDestroy(a->pul);
    MetricsDestroy(a->met);
    void* bufs[] = {a->dTraces, a->dBaselined, a->dStats,  a->dAgg,    a->dHist,      a->dEdge,
                    a->dModels, a->dLabels,    a->dLatent, a->dScore,  a->dPulses,    a->dCounts,
                    a->dOvf,    a->dPulseBl,   a->dIters,  a->dMetrics};
    for (void* p : bufs) DeviceFree(p);
    LearnedWsFree(a->lws);
    DeviceFree(a->dDiag);
    for (auto& e : a->ev)
        if (e) (void)hipEventDestroy(e);
    DeviceFree(a->dTraces2);
    for (int i = 0; i < 2; ++i) {
        if (a->copied[i]) (void)hipEventDestroy(a->copied[i]);
        if (a->bufFree[i]) (void)hipEventDestroy(a->bufFree[i]);
    }
    if (a->copyStream) (void)hipStreamDestroy(a->copyStream);
    if (a->stream) (void)hipStreamDestroy(a->stream);
    delete a;
}

#define HWPA_STAGE_BEGIN(a, k)                                          \
    do {                                                                \
        if ((a)->profiling) HWPA_HIP(hipEventRecord((a)->ev[k], (a)->stream)); \
        (a)->stageRan[k] = true;                                        \
    } while (0)

// DetectionModelEstimator::AddBatch (traceAnalysis/DetectionModelEstimator.cpp:93-160) on the
// chunk just baselined: aggregate statistics, histogram, model init before sequencing, and
// every MinFramesForEstimate frames an estimate followed by the histogram / aggregate reset.
// Sets a->ranEstimate when an estimate ran.
static int32_t DmeAddBatch(hwpa_analyzer* a, int32_t first_frame)
{
    const uint32_t L = a->lanes, F = a->frames;
    hipStream_t s = a->stream;
    int32_t rc;
    HWPA_STAGE_BEGIN(a, kStAggregate);
    a->plan |= kPlanAggregate;
    if ((rc = StatsMerge(a->dAgg, a->dStats, L, s)) != HWPA_OK) return rc;
    a->aggFi.Hull(first_frame, first_frame + (int32_t)F);
    if (a->status != kHistInit) {
        HWPA_STAGE_BEGIN(a, kStHistogram);
        a->plan |= kPlanHistAdd;
        if ((rc = HistAdd(a->dHist, a->dEdge, a->dBaselined, F, a->dModels, L, s, false)) != HWPA_OK) return rc;
        a->histFi.Hull(first_frame, first_frame + (int32_t)F);
    }
    HWPA_STAGE_BEGIN(a, kStDme);
    // The wrapping pass of the histogram (the lanes whose counts could pass 65535) opens this
    // stage, so the histogram stage's events time its fast pass alone.
    if (a->status != kHistInit) {
        if ((rc = HistAddWrap(a->dHist, a->dEdge, a->dBaselined, F, a->dModels, L, s)) != HWPA_OK) return rc;
    }
    if (a->status != kSequencing) {
        a->plan |= kPlanDmeInit;
        if ((rc = DmeInitModels(a->cfg, a->dAgg, a->dModels, L, s)) != HWPA_OK) return rc;
        a->modelFi = a->aggFi;
    }
    a->framesRemaining -= (int32_t)F;
    if (a->framesRemaining <= 0) {
        if (a->status == kDmeInit) a->status = kSequencing;
        if (a->status == kSequencing) {
            const int32_t mfi[2] = {a->modelFi.lo, a->modelFi.hi};
            const int32_t efi[2] = {a->histFi.lo, a->histFi.hi};
            if (a->mlp && a->cfg.learned_init) {
                rc = DmeEstimateLearned(a->cfg, a->mlp, a->lws, a->dHist, a->dAgg, a->dModels, mfi, efi, L, a->dIters, s);
                a->ranLearned = true;
                a->plan |= kPlanLearned;
            } else {
                const bool cap = a->capX != nullptr;
                rc = DmeEstimate(a->cfg, a->dHist, a->dAgg, a->dModels, mfi, efi, L, a->dIters, s,
                                 cap ? a->dDiag : nullptr);
                if (rc == HWPA_OK && cap) rc = MlpFeatures(a->dHist, L, a->capX, s);
                if (rc == HWPA_OK && cap)
                    rc = MlpTargets(a->dHist, a->dDiag, L, a->capMinConf, a->capT, a->capValid, s);
                a->ranLearned = false;
                a->plan |= kPlanEstimate | (cap ? kPlanCapture : 0u);
            }
            if (rc != HWPA_OK) return rc;
            a->modelFi = a->histFi;
            a->flags[1] = 1;
            a->ranEstimate = true;
            a->fullEstimation = true;
        }
        a->plan |= kPlanReset;
        // The next chunk's add always comes before the counters are read again, so
        // their zeros are left to it (chunks of up to 16384 frames).
        rc = (F <= 16384u) ? HistResetDeferred(a->cfg, a->dHist, a->dEdge, a->dAgg, L, s)
                           : HistReset(a->cfg, a->dHist, a->dAgg, L, s);
        if (rc != HWPA_OK) return rc;
        a->histFi = Interval{};
        if ((rc = StatsReset(a->dAgg, L, s)) != HWPA_OK) return rc;
        a->aggFi = Interval{};
        if (a->status == kHistInit) a->status = kDmeInit;
        a->framesRemaining = (int32_t)a->cfg.min_frames_for_estimate;
    }

    return HWPA_OK;
}

extern "C" {

int32_t hwpa_analyzer_create(uint32_t pool_id, uint32_t lanes, uint32_t frames, const hwpa_config* cfg,
                             hwpa_analyzer** out)
{
    if (!cfg || !out || lanes == 0) return Fail(HWPA_ERR_INVALID_ARG, "analyzer_create: bad argument");
    *out = nullptr;
    int32_t rc = hwpa_config_validate(cfg);
    if (rc != HWPA_OK) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return Fail(HWPA_ERR_NO_DEVICE, "no HIP device");
    const hwpa_filter_params fp = FilterParams(cfg->filter_type);
    if (frames < (uint32_t)NL || frames % (uint32_t)fp.aggregate_stride != 0)
        return Fail(HWPA_ERR_INVALID_ARG, "frames_per_chunk must be >= 24 and a multiple of the filter's aggregate stride");
    auto* a = new hwpa_analyzer;
    a->cfg = *cfg;
    a->poolId = pool_id;
    a->lanes = lanes;
    a->frames = frames;
    a->baselinerLatency = BaselinerStartupLatency(*cfg);
    const uint32_t cpe = (std::max(cfg->min_frames_for_estimate, 1u) + frames - 1) / frames;
    a->dmeDelay = (pool_id % cpe) * frames;
    a->framesRemaining = (int32_t)cfg->num_frames_pre_accum_stats;
    if (hipStreamCreate(&a->stream) != hipSuccess) {
        delete a;
        return Fail(HWPA_ERR_HIP, "hipStreamCreate failed");
    }
    for (auto& e : a->ev) {
        if (hipEventCreate(&e) != hipSuccess) {
            FreeAnalyzer(a);
            return Fail(HWPA_ERR_HIP, "hipEventCreate failed");
        }
    }
    const size_t trace = (size_t)lanes * frames * LS;
    const uint32_t mc = cfg->max_calls_per_zmw;
    struct A {
        void** p;
        size_t bytes;
        const char* what;
    } allocs[] = {
        {(void**)&a->dTraces, trace * 2, "traces"},
        {(void**)&a->dBaselined, trace * 2, "baselined"},
        {(void**)&a->dStats, sizeof(hwpa_baseliner_stats) * lanes, "stats"},
        {(void**)&a->dAgg, sizeof(hwpa_baseliner_stats) * lanes, "aggregate"},
        {(void**)&a->dHist, sizeof(hwpa_lane_hist) * lanes, "histograms"},
        {(void**)&a->dEdge, EdgeStateBytes() * lanes, "edge state"},
        {(void**)&a->dModels, sizeof(hwpa_lane_model) * lanes, "models"},
        {(void**)&a->dLabels, trace * 2, "labels"},
        {(void**)&a->dLatent, (size_t)lanes * NL * LS * 2, "latent"},
        {(void**)&a->dScore, (size_t)lanes * LS * 4, "score"},
        {(void**)&a->dPulses, (size_t)lanes * LS * mc * sizeof(hwpa_pulse), "pulses"},
        {(void**)&a->dCounts, (size_t)lanes * LS * 4, "counts"},
        {(void**)&a->dOvf, (size_t)lanes * LS * 4, "overflows"},
        {(void**)&a->dPulseBl, sizeof(hwpa_stat_acc) * lanes, "pulse stats"},
        {(void**)&a->dMetrics, sizeof(hwpa_basecalling_metrics) * lanes, "basecalling metrics"},
        {(void**)&a->dIters, sizeof(int32_t) * lanes, "EM iterations"},
    };
    for (auto& x : allocs) {
        rc = DeviceAlloc(x.p, x.bytes, x.what);
        if (rc != HWPA_OK) {
            FreeAnalyzer(a);
            return rc;
        }
    }
    if ((rc = BaselinerCreate(*cfg, lanes, &a->bl)) != HWPA_OK || (rc = LabelerCreate(*cfg, lanes, frames, &a->lab)) != HWPA_OK ||
        (rc = PulserCreate(*cfg, lanes, &a->pul)) != HWPA_OK || (rc = MetricsCreate(*cfg, lanes, &a->met)) != HWPA_OK) {
        FreeAnalyzer(a);
        return rc;
    }
    if ((rc = StatsReset(a->dAgg, lanes, a->stream)) != HWPA_OK || (rc = EdgeStateInit(a->dEdge, lanes, a->stream)) != HWPA_OK ||
        (rc = Fill32(a->dModels, 0, sizeof(hwpa_lane_model) * lanes / 4, a->stream)) != HWPA_OK ||
        (rc = Fill32(a->dCounts, 0, (size_t)lanes * LS, a->stream)) != HWPA_OK ||
        (rc = Fill32(a->dOvf, 0, (size_t)lanes * LS, a->stream)) != HWPA_OK ||
        (rc = Fill32(a->dIters, 0, lanes, a->stream)) != HWPA_OK) {
        FreeAnalyzer(a);
        return rc;
    }
    if (cfg->model_estimation_mode == HWPA_MODEL_FIXED) {
        std::vector<hwpa_lane_model> h(lanes);
        StaticModel(*cfg, &h[0]);
        for (uint32_t l = 1; l < lanes; ++l) h[l] = h[0];
        if (hipMemcpyAsync(a->dModels, h.data(), sizeof(hwpa_lane_model) * lanes, hipMemcpyHostToDevice, a->stream) !=
                hipSuccess ||
            hipStreamSynchronize(a->stream) != hipSuccess) {
            FreeAnalyzer(a);
            return Fail(HWPA_ERR_HIP, "analyzer: static model upload failed");
        }
    }
    if (hipStreamSynchronize(a->stream) != hipSuccess) {
        FreeAnalyzer(a);
        return Fail(HWPA_ERR_HIP, "analyzer init failed");
    }
    *out = a;
    return HWPA_OK;
}

int32_t hwpa_analyzer_destroy(hwpa_analyzer* a)
{
    FreeAnalyzer(a);
    return HWPA_OK;
}


// The host-input copy path, set up on the first chunk handed over in host memory: the copy
// stream (default priority), the second staging buffer and the events that order them.
static int32_t EnsureCopyPath(hwpa_analyzer* a)
{
    if (a->copyStream) return HWPA_OK;
    const size_t bytes = (size_t)a->lanes * a->frames * LS * 2;
    int32_t rc = DeviceAlloc((void**)&a->dTraces2, bytes, "traces (second staging buffer)");
    if (rc != HWPA_OK) return rc;
    for (int i = 0; i < 2; ++i) {
        if (hipEventCreateWithFlags(&a->copied[i], hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&a->bufFree[i], hipEventDisableTiming) != hipSuccess)
            return Fail(HWPA_ERR_HIP, "hipEventCreate failed");
    }
    if (hipStreamCreateWithFlags(&a->copyStream, hipStreamNonBlocking) != hipSuccess) {
        a->copyStream = nullptr;
        return Fail(HWPA_ERR_HIP, "hipStreamCreate failed");
    }
    return HWPA_OK;
}

// BatchAnalyzer::operator()(const TraceBatchVariant&) (analyzer/BatchAnalyzer.cpp:149-170) and
// the AnalyzeImpl of the configured mode: FixedModel (:172-205), SingleEstimate (:207-283),
// DynamicEstimate (:288-373).
static int32_t AnalyzerProcessImpl(hwpa_analyzer* a, const void* traces, int32_t encoding, int32_t on_device,
                                   uint32_t pool_id, int32_t first_frame)
{
    if (!a || !traces) return Fail(HWPA_ERR_INVALID_ARG, "analyzer_process: null argument");
    if (encoding != HWPA_ENCODING_INT16 && encoding != HWPA_ENCODING_UINT8)
        return Fail(HWPA_ERR_INVALID_ARG, "Unexpected data encoding in analyzer input");
    // BatchAnalyzer::operator() checks (BatchAnalyzer.cpp:151-152)
    if (pool_id != a->poolId) return Fail(HWPA_ERR_BAD_POOL, "Bad pool ID.");
    if (first_frame != a->nextFrame) return Fail(HWPA_ERR_BAD_FRAME, "Bad frame ID.");
    const uint32_t L = a->lanes, F = a->frames;
    const size_t traceBytes = (size_t)L * F * LS * (encoding == HWPA_ENCODING_UINT8 ? 1 : 2);
    hipStream_t s = a->stream;
    int32_t rc;
    std::memset(a->flags, 0, sizeof(a->flags));
    std::fill(std::begin(a->stageRan), std::end(a->stageRan), false);
    a->ranEstimate = false;
    a->plan = a->profiling ? kPlanEvents : 0u;
    if (encoding == HWPA_ENCODING_UINT8) a->plan |= kPlanUint8;

    const void* src = traces;
    int buf = -1;
    if (!on_device) {
        // Copy into the staging buffer the chunk before last used, once its baseline pass is
        // done, on the copy stream, while the previous chunk may still be in flight on `s`;
        // the chunk's stages are enqueued behind the copy, and the call returns once the copy
        // is complete (the caller's buffer is free again: AnalyzerProcess).
        if ((rc = EnsureCopyPath(a)) != HWPA_OK) return rc;
        buf = a->copyBuf;
        a->copyBuf ^= 1;
        int16_t* dst = buf ? a->dTraces2 : a->dTraces;
        if (a->bufUsed[buf]) HWPA_HIP(hipStreamWaitEvent(a->copyStream, a->bufFree[buf], 0));
        HWPA_HIP(hipMemcpyAsync(dst, traces, traceBytes, hipMemcpyHostToDevice, a->copyStream));
        HWPA_HIP(hipEventRecord(a->copied[buf], a->copyStream));
        a->pendingCopy = buf;
        HWPA_HIP(hipStreamWaitEvent(s, a->copied[buf], 0));
        src = dst;
        a->plan |= kPlanCopyIn;
    }
    HWPA_STAGE_BEGIN(a, kStBaseline);
    if ((rc = BaselinerRun(a->bl, src, F, a->dBaselined, a->dStats, s, encoding)) != HWPA_OK) return rc;
    if (buf >= 0) {
        HWPA_HIP(hipEventRecord(a->bufFree[buf], s));
        a->bufUsed[buf] = true;
    }
    a->nextFrame = first_frame + (int32_t)F;

    auto finish = [&](int32_t r) -> int32_t {
        if (a->profiling) {
            const hipError_t e = hipEventRecord(a->ev[kNumStages], s);
            if (e != hipSuccess) return Fail(HWPA_ERR_HIP, "hipEventRecord failed");
        }
        return r;
    };
    const size_t nz = (size_t)L * LS;
    auto zeroCounts = [&]() -> int32_t {
        hipLaunchKernelGGL(ZeroCountsKernel, dim3((unsigned)((nz + 255) / 256)), dim3(256), 0, s, a->dCounts, a->dOvf, nz);
        return CheckLaunch("ZeroCountsKernel");
    };
    auto runMetrics = [&]() -> int32_t {
        // HF metrics of this pulse batch, which ends NL latent frames before the chunk.
        HWPA_STAGE_BEGIN(a, kStMetrics);
        int32_t produced = 0;
        const int32_t r = MetricsProcess(a->met, a->dPulses, a->dCounts, a->cfg.max_calls_per_zmw, F,
                                         first_frame + (int32_t)F - NL, a->dStats, a->dScore, a->dPulseBl,
                                         a->dModels, a->dMetrics, &produced, s);
        if (r != HWPA_OK) return r;
        a->flags[3] = produced;
        a->plan |= kPlanMetrics;  // the finalize kernel is enqueued either way (same graph topology)
        return HWPA_OK;
    };
    auto labelsAndPulses = [&]() -> int32_t {
        HWPA_STAGE_BEGIN(a, kStLabels);
        a->plan |= kPlanLabels | kPlanPulses;
        int32_t r = LabelerRun(a->lab, a->dBaselined, F, a->dModels, a->dLabels, a->dLatent, a->dScore, s);
        if (r != HWPA_OK) return r;
        a->flags[2] = 1;
        HWPA_STAGE_BEGIN(a, kStPulses);
        return PulserRun(a->pul, a->dLabels, a->dLatent, a->dBaselined, F, first_frame - NL, a->dModels, a->dPulses,
                         a->dCounts, a->dOvf, a->dPulseBl, s);
    };

    if (a->cfg.model_estimation_mode == HWPA_MODEL_FIXED) {
        // FixedModelBatchAnalyzer: the static model, every stage every chunk.
        if ((rc = labelsAndPulses()) != HWPA_OK) return rc;
        a->flags[0] = 1;
        if ((rc = runMetrics()) != HWPA_OK) return rc;
        return finish(HWPA_OK);
    }

    if (a->cfg.model_estimation_mode == HWPA_MODEL_INITIAL) {
        // SingleEstimateBatchAnalyzer: feed the DME after the baseliner start-up until its first
        // estimate, then keep that model; empty pulses (and empty labeler / pulse-detector
        // metrics: score 0, no baseline frames) until then.  Metrics run every chunk.
        if (!a->modelInitialized && first_frame > (int32_t)a->baselinerLatency) {
            if ((rc = DmeAddBatch(a, first_frame)) != HWPA_OK) return rc;
            a->modelInitialized = a->ranEstimate;
        }
        if (a->modelInitialized) {
            if ((rc = labelsAndPulses()) != HWPA_OK) return rc;
            a->flags[0] = 1;
        } else {
            a->plan |= kPlanEmptyPulses;
            if ((rc = zeroCounts()) != HWPA_OK) return rc;
            if ((rc = Fill32(a->dScore, 0, nz, s)) != HWPA_OK) return rc;
            if ((rc = Fill32(a->dPulseBl, 0, sizeof(hwpa_stat_acc) * L / 4, s)) != HWPA_OK) return rc;
        }
        if ((rc = runMetrics()) != HWPA_OK) return rc;
        return finish(HWPA_OK);
    }

    // DynamicEstimateBatchAnalyzer.  Early empty result during the baseliner start-up and the
    // pool stagger (BatchAnalyzer.cpp:331-337).
    if (first_frame < (int32_t)(a->baselinerLatency + a->dmeDelay)) {
        a->plan |= kPlanEarlyZero;
        if ((rc = zeroCounts()) != HWPA_OK) return rc;
        return finish(HWPA_OK);
    }
    if ((rc = DmeAddBatch(a, first_frame)) != HWPA_OK) return rc;
    if ((rc = labelsAndPulses()) != HWPA_OK) return rc;
    if (a->fullEstimation) {
        a->flags[0] = 1;
        if ((rc = runMetrics()) != HWPA_OK) return rc;
    } else {
        a->plan |= kPlanLateZero;
        if ((rc = zeroCounts()) != HWPA_OK) return rc;
    }
    return finish(HWPA_OK);
}

// The host-input copy of this call (if any) is waited for on every return path, errors included.
static int32_t AnalyzerProcess(hwpa_analyzer* a, const void* traces, int32_t encoding, int32_t on_device,
                               uint32_t pool_id, int32_t first_frame)
{
    const int32_t rc = AnalyzerProcessImpl(a, traces, encoding, on_device, pool_id, first_frame);
    if (a && a->pendingCopy >= 0) {
        const int b = a->pendingCopy;
        a->pendingCopy = -1;
        if (hipEventSynchronize(a->copied[b]) != hipSuccess && rc == HWPA_OK)
            return Fail(HWPA_ERR_HIP, "host chunk copy failed");
    }
    return rc;
}

int32_t hwpa_analyzer_process(hwpa_analyzer* a, const int16_t* traces, int32_t on_device, uint32_t pool_id,
                              int32_t first_frame)
{
    return AnalyzerProcess(a, traces, HWPA_ENCODING_INT16, on_device, pool_id, first_frame);
}

int32_t hwpa_analyzer_process_encoded(hwpa_analyzer* a, const void* traces, int32_t encoding, int32_t on_device,
                                      uint32_t pool_id, int32_t first_frame)
{
    return AnalyzerProcess(a, traces, encoding, on_device, pool_id, first_frame);
}

int32_t hwpa_analyzer_sync(hwpa_analyzer* a)
{
    if (!a) return Fail(HWPA_ERR_INVALID_ARG, "null analyzer");
    HWPA_HIP(hipStreamSynchronize(a->stream));
    return HWPA_OK;
}

int32_t hwpa_result_flags(const hwpa_analyzer* a, int32_t flags[4])
{
    if (!a || !flags) return Fail(HWPA_ERR_INVALID_ARG, "null argument");
    std::memcpy(flags, a->flags, sizeof(a->flags));
    return HWPA_OK;
}

int32_t hwpa_result_pulses(const hwpa_analyzer* a, const hwpa_pulse** d_pulses, const uint32_t** d_counts,
                           const uint32_t** d_overflows, uint32_t* max_calls)
{
    if (!a) return Fail(HWPA_ERR_INVALID_ARG, "null analyzer");
    if (d_pulses) *d_pulses = a->dPulses;
    if (d_counts) *d_counts = a->dCounts;
    if (d_overflows) *d_overflows = a->dOvf;
    if (max_calls) *max_calls = a->cfg.max_calls_per_zmw;
    return HWPA_OK;
}

int32_t hwpa_result_labels(const hwpa_analyzer* a, const int16_t** d)
{
    if (!a || !d) return Fail(HWPA_ERR_INVALID_ARG, "null argument");
    *d = a->dLabels;
    return HWPA_OK;
}

int32_t hwpa_result_baselined(const hwpa_analyzer* a, const int16_t** d)
{
    if (!a || !d) return Fail(HWPA_ERR_INVALID_ARG, "null argument");
    *d = a->dBaselined;
    return HWPA_OK;
}

int32_t hwpa_result_models(const hwpa_analyzer* a, const hwpa_lane_model** d)
{
    if (!a || !d) return Fail(HWPA_ERR_INVALID_ARG, "null argument");
    *d = a->dModels;
    return HWPA_OK;
}

int32_t hwpa_result_baseliner_stats(const hwpa_analyzer* a, const hwpa_baseliner_stats** d)
{
    if (!a || !d) return Fail(HWPA_ERR_INVALID_ARG, "null argument");
    *d = a->dStats;
    return HWPA_OK;
}

int32_t hwpa_result_pulse_baseline_stats(const hwpa_analyzer* a, const hwpa_stat_acc** d)
{
    if (!a || !d) return Fail(HWPA_ERR_INVALID_ARG, "null argument");
    *d = a->dPulseBl;
    return HWPA_OK;
}

int32_t hwpa_result_viterbi_score(const hwpa_analyzer* a, const float** d)
{
    if (!a || !d) return Fail(HWPA_ERR_INVALID_ARG, "null argument");
    *d = a->dScore;
    return HWPA_OK;
}

int32_t hwpa_result_metrics(const hwpa_analyzer* a, const hwpa_basecalling_metrics** d)
{
    if (!a || !d) return Fail(HWPA_ERR_INVALID_ARG, "null argument");
    *d = a->dMetrics;
    return HWPA_OK;
}

int32_t hwpa_analyzer_set_activity_tree(hwpa_analyzer* a, const hwpa_activity_tree* tree)
{
    if (!a) return Fail(HWPA_ERR_INVALID_ARG, "null analyzer");
    return MetricsSetTree(a->met, tree);
}

int32_t hwpa_analyzer_set_profiling(hwpa_analyzer* a, int32_t enable)
{
    if (!a) return Fail(HWPA_ERR_INVALID_ARG, "null analyzer");
    a->profiling = enable != 0;
    return HWPA_OK;
}

int32_t hwpa_analyzer_stage_ms(const hwpa_analyzer* a, float ms[8])
{
    if (!a || !ms) return Fail(HWPA_ERR_INVALID_ARG, "null argument");
    for (int i = 0; i < 8; ++i) ms[i] = 0.0f;
    if (!a->profiling) return Fail(HWPA_ERR_INVALID_ARG, "profiling is not enabled");
    HWPA_HIP(hipEventSynchronize(a->ev[kNumStages]));
    // Each stage runs from its begin event to the next recorded event.
    for (int k = 0; k < kNumStages; ++k) {
        if (!a->stageRan[k]) continue;
        int next = kNumStages;
        for (int j = k + 1; j < kNumStages; ++j)
            if (a->stageRan[j]) {
                next = j;
                break;
            }
        float t = 0.0f;
        HWPA_HIP(hipEventElapsedTime(&t, a->ev[k], a->ev[next]));
        ms[k] = t;
    }
    return HWPA_OK;
}

int32_t hwpa_analyzer_em_iterations(const hwpa_analyzer* a, int32_t* iters)
{
    if (!a || !iters) return Fail(HWPA_ERR_INVALID_ARG, "null argument");
    *iters = 0;
    if (!a->ranEstimate) return HWPA_OK;
    std::vector<int32_t> h(a->lanes);
    HWPA_HIP(hipStreamSynchronize(a->stream));
    HWPA_HIP(hipMemcpy(h.data(), a->dIters, sizeof(int32_t) * a->lanes, hipMemcpyDeviceToHost));
    for (int32_t v : h) *iters = std::max(*iters, v);
    return HWPA_OK;
}

int32_t hwpa_analyzer_set_initializer(hwpa_analyzer* a, hwpa_initializer* m)
{
    if (!a) return Fail(HWPA_ERR_INVALID_ARG, "null analyzer");
    HWPA_HIP(hipStreamSynchronize(a->stream));
    if (m && a->lws.lanes < a->lanes) {
        const int32_t rc = LearnedWsAlloc(a->lws, a->lanes);
        if (rc != HWPA_OK) return rc;
    }
    a->mlp = m;
    return HWPA_OK;
}

int32_t hwpa_analyzer_capture_training(hwpa_analyzer* a, uint16_t* d_x, float* d_t, uint8_t* d_valid,
                                       float min_confidence)
{
    if (!a) return Fail(HWPA_ERR_INVALID_ARG, "null analyzer");
    if (d_x && (!d_t || !d_valid)) return Fail(HWPA_ERR_INVALID_ARG, "capture: targets and valid flags required");
    if (d_x && ((uintptr_t)d_x & 15) != 0) return Fail(HWPA_ERR_INVALID_ARG, "capture: input must be 16-byte aligned");
    HWPA_HIP(hipStreamSynchronize(a->stream));
    if (d_x && !a->dDiag) {
        const int32_t rc = DeviceAlloc((void**)&a->dDiag, (size_t)a->lanes * HWPA_EM_DIAG_ROWS * LS * sizeof(float),
                                       "EM diagnostics");
        if (rc != HWPA_OK) return rc;
    }
    a->capX = d_x;
    a->capT = d_x ? d_t : nullptr;
    a->capValid = d_x ? d_valid : nullptr;
    a->capMinConf = min_confidence;
    return HWPA_OK;
}

int32_t hwpa_analyzer_fallback_count(const hwpa_analyzer* a, uint32_t* count)
{
    if (!a || !count) return Fail(HWPA_ERR_INVALID_ARG, "null argument");
    *count = 0;
    if (!a->ranEstimate || !a->ranLearned) return HWPA_OK;
    HWPA_HIP(hipStreamSynchronize(a->stream));
    HWPA_HIP(hipMemcpy(count, a->lws.total + 1, sizeof(uint32_t), hipMemcpyDeviceToHost));
    return HWPA_OK;
}

int32_t hwpa_analyzer_stream(const hwpa_analyzer* a, void** stream)
{
    if (!a || !stream) return Fail(HWPA_ERR_INVALID_ARG, "null argument");
    *stream = (void*)a->stream;
    return HWPA_OK;
}

}  // extern "C"

// ---- batched pools: one captured graph per chunk step ----------------------
// Every pool keeps its own analyzer (state machine, buffers, stream).  A step
// captures all pools' work from a master stream that forks to the pool
// streams and joins them back, so the whole step is one graph launch with the
// pools running concurrently.  The analyzers' host state machines run during
// capture, exactly as for direct calls; the captured graph is launched at once.
// Executable graphs are cached by the pools' plans (the kernel sequence each
// pool enqueued) and refreshed with hipGraphExecUpdate, which only rewrites
// kernel arguments (chunk pointers, frame intervals) when the topology matches.

struct hwpa_batch {
    std::vector<hwpa_analyzer*> pools;
    hipStream_t master = nullptr;
    hipEvent_t fork = nullptr, done = nullptr;
    std::vector<hipEvent_t> joins;
    bool useGraph = true;
    struct Exec {
        std::vector<uint32_t> key;
        hipGraphExec_t exec = nullptr;
    };
    std::vector<Exec> cache;
    uint64_t steps = 0, updates = 0, instantiations = 0;
};

static void FreeBatch(hwpa_batch* b)
{
    if (!b) return;
    if (b->master) (void)hipStreamSynchronize(b->master);
    for (auto& e : b->cache)
        if (e.exec) (void)hipGraphExecDestroy(e.exec);
    for (auto* a : b->pools) FreeAnalyzer(a);
    for (auto& e : b->joins)
        if (e) (void)hipEventDestroy(e);
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->done) (void)hipEventDestroy(b->done);
    if (b->master) (void)hipStreamDestroy(b->master);
    delete b;
}

extern "C" {

int32_t hwpa_batch_create(uint32_t n_pools, const uint32_t* pool_ids, const uint32_t* lanes, const uint32_t* frames,
                          const hwpa_config* cfgs, hwpa_batch** out)
{
    if (!out || !pool_ids || !lanes || !frames || !cfgs || n_pools == 0)
        return Fail(HWPA_ERR_INVALID_ARG, "batch_create: bad argument");
    *out = nullptr;
    hwpa_batch* b = new (std::nothrow) hwpa_batch();
    if (!b) return Fail(HWPA_ERR_OOM, "out of host memory");
    for (uint32_t i = 0; i < n_pools; ++i) {
        hwpa_analyzer* a = nullptr;
        const int32_t rc = hwpa_analyzer_create(pool_ids[i], lanes[i], frames[i], &cfgs[i], &a);
        if (rc != HWPA_OK) {
            FreeBatch(b);
            return rc;
        }
        b->pools.push_back(a);
    }
    b->joins.assign(n_pools, nullptr);
    bool ok = hipStreamCreate(&b->master) == hipSuccess;
    ok = ok && hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) == hipSuccess;
    ok = ok && hipEventCreateWithFlags(&b->done, hipEventDisableTiming) == hipSuccess;
    for (auto& e : b->joins) ok = ok && hipEventCreateWithFlags(&e, hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        FreeBatch(b);
        return Fail(HWPA_ERR_HIP, "batch: stream/event creation failed");
    }
    *out = b;
    return HWPA_OK;
}

int32_t hwpa_batch_destroy(hwpa_batch* b)
{
    FreeBatch(b);
    return HWPA_OK;
}

int32_t hwpa_batch_pool(hwpa_batch* b, uint32_t index, hwpa_analyzer** a)
{
    if (!b || !a || index >= b->pools.size()) return Fail(HWPA_ERR_INVALID_ARG, "batch_pool: bad argument");
    *a = b->pools[index];
    return HWPA_OK;
}

int32_t hwpa_batch_set_graph(hwpa_batch* b, int32_t enable)
{
    if (!b) return Fail(HWPA_ERR_INVALID_ARG, "null batch");
    b->useGraph = enable != 0;
    return HWPA_OK;
}

int32_t hwpa_batch_set_initializer(hwpa_batch* b, hwpa_initializer* m)
{
    if (!b) return Fail(HWPA_ERR_INVALID_ARG, "null batch");
    for (auto* a : b->pools) {
        const int32_t rc = hwpa_analyzer_set_initializer(a, m);
        if (rc != HWPA_OK) return rc;
    }
    return HWPA_OK;
}

static int32_t BatchProcess(hwpa_batch* b, const void* const* d_traces, const int32_t* encodings,
                            const int32_t* first_frames)
{
    if (!b || !d_traces || !first_frames) return Fail(HWPA_ERR_INVALID_ARG, "batch_process: null argument");
    const size_t n = b->pools.size();
    auto enc = [&](size_t i) { return encodings ? encodings[i] : (int32_t)HWPA_ENCODING_INT16; };
    if (!b->useGraph) {
        for (size_t i = 0; i < n; ++i) {
            hwpa_analyzer* a = b->pools[i];
            const int32_t rc = AnalyzerProcess(a, d_traces[i], enc(i), 1, a->poolId, first_frames[i]);
            if (rc != HWPA_OK) return rc;
        }
        b->steps += 1;
        return HWPA_OK;
    }
    for (auto* a : b->pools)
        if (a->profiling) return Fail(HWPA_ERR_UNSUPPORTED, "batch: stage profiling is not captured into graphs");
    HWPA_HIP(hipStreamBeginCapture(b->master, hipStreamCaptureModeThreadLocal));
    int32_t rc = HWPA_OK;
    if (hipEventRecord(b->fork, b->master) != hipSuccess) rc = Fail(HWPA_ERR_HIP, "batch: fork record failed");
    for (size_t i = 0; i < n && rc == HWPA_OK; ++i) {
        hwpa_analyzer* a = b->pools[i];
        if (hipStreamWaitEvent(a->stream, b->fork, 0) != hipSuccess) {
            rc = Fail(HWPA_ERR_HIP, "batch: fork wait failed");
            break;
        }
        rc = AnalyzerProcess(a, d_traces[i], enc(i), 1, a->poolId, first_frames[i]);
        if (rc != HWPA_OK) break;
        if (hipEventRecord(b->joins[i], a->stream) != hipSuccess ||
            hipStreamWaitEvent(b->master, b->joins[i], 0) != hipSuccess)
            rc = Fail(HWPA_ERR_HIP, "batch: join failed");
    }
    hipGraph_t g = nullptr;
    const hipError_t ce = hipStreamEndCapture(b->master, &g);
    if (rc != HWPA_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (ce != hipSuccess || !g) return Fail(HWPA_ERR_HIP, "batch: stream capture failed");
    std::vector<uint32_t> key(n);
    for (size_t i = 0; i < n; ++i) key[i] = b->pools[i]->plan;
    hipGraphExec_t exec = nullptr;
    for (auto& e : b->cache) {
        if (e.key != key) continue;
        hipGraphNode_t errNode = nullptr;
        hipGraphExecUpdateResult res;
        if (hipGraphExecUpdate(e.exec, g, &errNode, &res) == hipSuccess && res == hipGraphExecUpdateSuccess) {
            exec = e.exec;
            b->updates += 1;
        } else {
            (void)hipGetLastError();
            (void)hipGraphExecDestroy(e.exec);
            e.exec = nullptr;
            if (hipGraphInstantiate(&e.exec, g, nullptr, nullptr, 0) != hipSuccess) {
                (void)hipGraphDestroy(g);
                return Fail(HWPA_ERR_HIP, "batch: graph instantiation failed");
            }
            exec = e.exec;
            b->instantiations += 1;
        }
        break;
    }
    if (!exec) {
        hwpa_batch::Exec e;
        e.key = key;
        if (hipGraphInstantiate(&e.exec, g, nullptr, nullptr, 0) != hipSuccess) {
            (void)hipGraphDestroy(g);
            return Fail(HWPA_ERR_HIP, "batch: graph instantiation failed");
        }
        exec = e.exec;
        b->cache.push_back(e);
        b->instantiations += 1;
    }
    const hipError_t le = hipGraphLaunch(exec, b->master);
    (void)hipGraphDestroy(g);
    if (le != hipSuccess) return Fail(HWPA_ERR_HIP, "batch: graph launch failed");
    // Let each pool's stream (and so hwpa_analyzer_sync / result accessors) see the step.
    HWPA_HIP(hipEventRecord(b->done, b->master));
    for (auto* a : b->pools) HWPA_HIP(hipStreamWaitEvent(a->stream, b->done, 0));
    b->steps += 1;
    return HWPA_OK;
}

int32_t hwpa_batch_process(hwpa_batch* b, const int16_t* const* d_traces, const int32_t* first_frames)
{
    if (!b || !d_traces) return Fail(HWPA_ERR_INVALID_ARG, "batch_process: null argument");
    std::vector<const void*> p(d_traces, d_traces + b->pools.size());
    return BatchProcess(b, p.data(), nullptr, first_frames);
}

int32_t hwpa_batch_process_encoded(hwpa_batch* b, const void* const* d_traces, const int32_t* encodings,
                                   const int32_t* first_frames)
{
    return BatchProcess(b, d_traces, encodings, first_frames);
}

int32_t hwpa_batch_sync(hwpa_batch* b)
{
    if (!b) return Fail(HWPA_ERR_INVALID_ARG, "null batch");
    HWPA_HIP(hipStreamSynchronize(b->master));
    for (auto* a : b->pools) HWPA_HIP(hipStreamSynchronize(a->stream));
    return HWPA_OK;
}

int32_t hwpa_batch_autotune(hwpa_batch* b, int32_t repeats)
{
    if (!b) return Fail(HWPA_ERR_INVALID_ARG, "null batch");
    HWPA_HIP(hipStreamSynchronize(b->master));
    for (size_t i = 0; i < b->pools.size(); ++i) {
        const hwpa_analyzer* a = b->pools[i];
        bool seen = false;
        for (size_t j = 0; j < i && !seen; ++j)
            seen = b->pools[j]->lanes == a->lanes && b->pools[j]->frames == a->frames;
        if (seen) continue;
        const int32_t rc = hwpa_autotune(&a->cfg, a->lanes, a->frames, repeats, nullptr, 0);
        if (rc != HWPA_OK) return rc;
    }
    return HWPA_OK;
}

int32_t hwpa_batch_graph_stats(const hwpa_batch* b, uint64_t stats[3])
{
    if (!b || !stats) return Fail(HWPA_ERR_INVALID_ARG, "null argument");
    stats[0] = b->steps;
    stats[1] = b->instantiations;
    stats[2] = b->updates;
    return HWPA_OK;
}

}  // extern "C"
