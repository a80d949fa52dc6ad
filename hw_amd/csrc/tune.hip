// tune.hip -- autotuner of kernel shapes.  No reference counterpart.
//
// Tunable kernels come in variants that compute the same bits with different
// shapes: the baseline filter's lane tile (lanes per workgroup), the
// histogram's frame block (frames per prefetched batch) and LDS bin window,
// and the frame batches of the Viterbi labeler and the pulse accumulator.
// The table maps (kernel, lanes, frames) to the variant to launch; the
// autotuner fills it by timing every variant on synthetic traces of the
// shape with HIP events, and the launchers read it on every call.

#include "internal.h"

#include <map>
#include <mutex>
#include <tuple>
#include <vector>

namespace hwpa {

static std::mutex g_tuneMutex;
static std::map<std::tuple<int, uint32_t, uint32_t>, int32_t> g_tune;

int32_t TuneVariants(int kernel)
{
    switch (kernel) {
    case kTuneHistogram: return 22;
    case kTuneBaseline: return 6;
    case kTuneLabel: return 4;
    case kTunePulse: return 13;
    default: return 0;
    }
}

int32_t TuneGet(int kernel, uint32_t lanes, uint32_t frames)
{
    std::lock_guard<std::mutex> lk(g_tuneMutex);
    const auto it = g_tune.find(std::make_tuple(kernel, lanes, frames));
    return it == g_tune.end() ? 0 : it->second;
}

void TuneSet(int kernel, uint32_t lanes, uint32_t frames, int32_t variant)
{
    std::lock_guard<std::mutex> lk(g_tuneMutex);
    g_tune[std::make_tuple(kernel, lanes, frames)] = variant;
}

namespace {

struct Timer {
    hipEvent_t a = nullptr, b = nullptr;
    bool ok = false;
    Timer() { ok = hipEventCreate(&a) == hipSuccess && hipEventCreate(&b) == hipSuccess; }
    ~Timer()
    {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

struct Scratch {
    std::vector<void*> ptrs;
    ~Scratch()
    {
        for (void* p : ptrs) DeviceFree(p);
    }
    template <typename T>
    int32_t Alloc(T** p, size_t bytes)
    {
        const int32_t rc = DeviceAlloc((void**)p, bytes, "autotune scratch");
        if (rc == HWPA_OK) ptrs.push_back((void*)*p);
        return rc;
    }
};

}  // namespace

// Times `run` (which enqueues one launch sequence on s) `repeats` times after
// one warm-up call; returns the mean milliseconds or a negative value.
template <class F>
static float TimeIt(hipStream_t s, int repeats, F run)
{
    Timer t;
    if (!t.ok) return -1.0f;
    if (run() != HWPA_OK) return -1.0f;
    if (hipEventRecord(t.a, s) != hipSuccess) return -1.0f;
    for (int r = 0; r < repeats; ++r)
        if (run() != HWPA_OK) return -1.0f;
    if (hipEventRecord(t.b, s) != hipSuccess || hipEventSynchronize(t.b) != hipSuccess) return -1.0f;
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, t.a, t.b) != hipSuccess) return -1.0f;
    return ms / (float)repeats;
}

}  // namespace hwpa

using namespace hwpa;

extern "C" {

int32_t hwpa_tune_variants(int32_t kernel, int32_t* count)
{
    if (!count) return Fail(HWPA_ERR_INVALID_ARG, "null argument");
    *count = TuneVariants(kernel);
    return *count > 0 ? HWPA_OK : Fail(HWPA_ERR_INVALID_ARG, "unknown tunable kernel");
}

int32_t hwpa_tune_set(int32_t kernel, uint32_t lanes, uint32_t frames, int32_t variant)
{
    if (variant < 0 || variant >= TuneVariants(kernel)) return Fail(HWPA_ERR_INVALID_ARG, "unknown kernel variant");
    TuneSet(kernel, lanes, frames, variant);
    return HWPA_OK;
}

int32_t hwpa_tune_get(int32_t kernel, uint32_t lanes, uint32_t frames, int32_t* variant)
{
    if (!variant || TuneVariants(kernel) == 0) return Fail(HWPA_ERR_INVALID_ARG, "bad argument");
    *variant = TuneGet(kernel, lanes, frames);
    return HWPA_OK;
}

int32_t hwpa_autotune(const hwpa_config* cfg, uint32_t lanes, uint32_t frames, int32_t repeats, float* ms,
                      int32_t ms_len)
{
    if (!cfg || lanes == 0 || frames == 0 || repeats <= 0) return Fail(HWPA_ERR_INVALID_ARG, "autotune: bad argument");
    int32_t rc = hwpa_config_validate(cfg);
    if (rc != HWPA_OK) return rc;
    hipStream_t s = nullptr;
    HWPA_HIP(hipStreamCreate(&s));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } guard{s};
    Scratch sc;
    const size_t trace = (size_t)lanes * frames * LS;
    int16_t *raw = nullptr, *bls = nullptr;
    hwpa_baseliner_stats* st = nullptr;
    hwpa_lane_hist* hist = nullptr;
    hwpa_lane_model* models = nullptr;
    hwpa_tracegen_state* gst = nullptr;
    void* edge = nullptr;
    if ((rc = sc.Alloc(&raw, trace * 2)) != HWPA_OK || (rc = sc.Alloc(&bls, trace * 2)) != HWPA_OK ||
        (rc = sc.Alloc(&st, lanes * sizeof(hwpa_baseliner_stats))) != HWPA_OK ||
        (rc = sc.Alloc(&hist, lanes * sizeof(hwpa_lane_hist))) != HWPA_OK ||
        (rc = sc.Alloc(&models, lanes * sizeof(hwpa_lane_model))) != HWPA_OK ||
        (rc = sc.Alloc(&gst, lanes * sizeof(hwpa_tracegen_state))) != HWPA_OK ||
        (rc = sc.Alloc(&edge, lanes * EdgeStateBytes())) != HWPA_OK)
        return rc;
    hwpa_tracegen_params gp;
    if ((rc = hwpa_tracegen_default(&gp)) != HWPA_OK) return rc;
    if ((rc = hwpa_tracegen_run(&gp, lanes, 0, frames, 0, gst, 1, raw, s)) != HWPA_OK) return rc;

    BaselinerDev* bl = nullptr;
    if ((rc = BaselinerCreate(*cfg, lanes, &bl)) != HWPA_OK) return rc;
    struct BlGuard {
        BaselinerDev* b;
        ~BlGuard() { BaselinerDestroy(b); }
    } blg{bl};
    // Labeler and pulser inputs: labels, latent frames, scores and pulse slots.
    const bool stages = frames > (uint32_t)NL;
    LabelerDev* lab = nullptr;
    PulserDev* pul = nullptr;
    if (stages) {
        if ((rc = LabelerCreate(*cfg, lanes, frames, &lab)) != HWPA_OK) return rc;
        if ((rc = PulserCreate(*cfg, lanes, &pul)) != HWPA_OK) {
            LabelerDestroy(lab);
            return rc;
        }
    }
    struct StageGuard {
        LabelerDev* l;
        PulserDev* p;
        ~StageGuard()
        {
            if (l) LabelerDestroy(l);
            if (p) PulserDestroy(p);
        }
    } sg{lab, pul};
    int16_t *labels = nullptr, *latent = nullptr;
    float* score = nullptr;
    hwpa_pulse* pulses = nullptr;
    uint32_t *counts = nullptr, *ovf = nullptr;
    hwpa_stat_acc* pbl = nullptr;
    if (stages) {
        const uint32_t mc = PulserMaxCalls(pul);
        if ((rc = sc.Alloc(&labels, trace * 2)) != HWPA_OK ||
            (rc = sc.Alloc(&latent, (size_t)lanes * NL * LS * 2)) != HWPA_OK ||
            (rc = sc.Alloc(&score, (size_t)lanes * LS * 4)) != HWPA_OK ||
            (rc = sc.Alloc(&pulses, (size_t)lanes * LS * mc * sizeof(hwpa_pulse))) != HWPA_OK ||
            (rc = sc.Alloc(&counts, (size_t)lanes * LS * 4)) != HWPA_OK ||
            (rc = sc.Alloc(&ovf, (size_t)lanes * LS * 4)) != HWPA_OK ||
            (rc = sc.Alloc(&pbl, lanes * sizeof(hwpa_stat_acc))) != HWPA_OK)
            return rc;
    }

    std::vector<float> times;
    for (int kernel = 0; kernel < kTuneKernels; ++kernel) {
        const int nv = TuneVariants(kernel);
        int best = 0;
        float bestMs = -1.0f;
        for (int v = 0; v < nv; ++v) {
            float t = -1.0f;
            if (kernel == kTuneBaseline) {
                TuneSet(kTuneBaseline, lanes, frames, v);
                t = TimeIt(s, repeats, [&]() { return BaselinerRun(bl, raw, frames, bls, st, s); });
            } else if (kernel == kTuneLabel || kernel == kTunePulse) {
                // Labels of the baselined traces under the start-up models.
                if (!stages) {
                    t = 0.0f;
                } else if (kernel == kTuneLabel) {
                    TuneSet(kTuneLabel, lanes, frames, v);
                    t = TimeIt(s, repeats, [&]() { return LabelerRun(lab, bls, frames, models, labels, latent, score, s); });
                } else {
                    TuneSet(kTunePulse, lanes, frames, v);
                    t = TimeIt(s, repeats, [&]() {
                        return PulserRun(pul, labels, latent, bls, frames, 0, models, pulses, counts, ovf, pbl, s);
                    });
                }
            } else {
                // Histogram inputs: this shape's baselined traces, bounds and models from its stats.
                if (v == 0) {
                    if ((rc = BaselinerRun(bl, raw, frames, bls, st, s)) != HWPA_OK) return rc;
                    if ((rc = DmeInitModels(*cfg, st, models, lanes, s)) != HWPA_OK) return rc;
                    if ((rc = EdgeStateInit(edge, lanes, s)) != HWPA_OK) return rc;
                }
                // As the analyzer runs it after an estimate: the reset leaves the zero
                // counters to the add.
                t = TimeIt(s, repeats, [&]() {
                    const int32_t r = HistResetDeferred(*cfg, hist, edge, st, lanes, s);
                    return r != HWPA_OK ? r : HistAddVariant(v, hist, edge, bls, frames, models, lanes, s);
                });
            }
            if (t < 0.0f) return Fail(HWPA_ERR_HIP, "autotune: timing failed");
            times.push_back(t);
            if (bestMs < 0.0f || t < bestMs) {
                bestMs = t;
                best = v;
            }
        }
        TuneSet(kernel, lanes, frames, best);
    }
    if (ms) {
        for (int32_t i = 0; i < ms_len; ++i) ms[i] = (i < (int32_t)times.size()) ? times[i] : 0.0f;
    }
    HWPA_HIP(hipStreamSynchronize(s));
    return HWPA_OK;
}

}  // extern "C"
