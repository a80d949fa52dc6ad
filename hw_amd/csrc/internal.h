// internal.h -- shared declarations of libhwpa (not part of the C-ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/hwpa.h"
#include "../../include/hwpa_math.h"

namespace hwpa {

constexpr int LS = HWPA_LANE_SIZE;
constexpr int NA = HWPA_NUM_ANALOGS;
constexpr int NB = HWPA_NUM_BINS;
constexpr int NS = HWPA_NUM_STATES;
constexpr int NL = HWPA_LATENT_FRAMES;

// Thread-local error text behind hwpa_last_error().
void SetError(const std::string& msg);
void ClearError();

int32_t Fail(int32_t code, const std::string& msg);

#define HWPA_HIP(expr)                                                                  \
    do {                                                                                \
        hipError_t hwpa_e_ = (expr);                                                    \
        if (hwpa_e_ != hipSuccess)                                                      \
            return ::hwpa::Fail(HWPA_ERR_HIP, std::string(#expr) + ": " +              \
                                                  hipGetErrorString(hwpa_e_));          \
    } while (0)

inline hipStream_t AsStream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Reports and clears any error left by the last kernel launch.
int32_t CheckLaunch(const char* what);

// Device allocation helpers (hipMalloc / hipFree with error text).
int32_t DeviceAlloc(void** p, size_t bytes, const char* what);
void DeviceFree(void* p);

// Fills `words` 32-bit words at d with `value` using a kernel (no memset
// nodes, so the call is safe inside captured graphs).
int32_t Fill32(void* d, uint32_t value, size_t words, hipStream_t s);

// Host-side constant derivations shared by several stages.
hwpa_filter_params FilterParams(int32_t type);

struct DmeConsts {
    float blVarMin, blVarMax, blVarNominal, signalScaler, shotVarCoeff;
};
DmeConsts MakeDmeConsts(const hwpa_config& c);

// ---- stage launchers implemented in the .hip files ----------------------

struct BaselinerDev;
int32_t BaselinerCreate(const hwpa_config& cfg, uint32_t lanes, BaselinerDev** out);
void BaselinerDestroy(BaselinerDev* b);
// raw: [lane][frame][64] samples of the given HWPA_ENCODING_* type.
int32_t BaselinerRun(BaselinerDev* b, const void* raw, uint32_t frames, int16_t* out,
                     hwpa_baseliner_stats* stats, hipStream_t s, int32_t encoding = HWPA_ENCODING_INT16);
uint32_t BaselinerStartupLatency(const hwpa_config& cfg);

int32_t StatsReset(hwpa_baseliner_stats* agg, uint32_t lanes, hipStream_t s);
int32_t StatsMerge(hwpa_baseliner_stats* agg, const hwpa_baseliner_stats* chunk, uint32_t lanes,
                   hipStream_t s);

size_t EdgeStateBytes();
int32_t EdgeStateInit(void* edge, uint32_t lanes, hipStream_t s);
int32_t HistReset(const hwpa_config& cfg, hwpa_lane_hist* hist, const hwpa_baseliner_stats* stats,
                  uint32_t lanes, hipStream_t s);
// HistReset that leaves the counters for the next add to start from zero (no
// counters written or read back); that add must precede any read of the counters.
int32_t HistResetDeferred(const hwpa_config& cfg, hwpa_lane_hist* hist, void* edge, const hwpa_baseliner_stats* stats,
                          uint32_t lanes, hipStream_t s);
// Fast pass over the chunk; with wrap, also the 16-bit wrapping pass over the lanes it left
// (lanes whose counts could pass 65535).  Without it, HistAddWrap must run before the
// traces change or the counters are read.
int32_t HistAdd(hwpa_lane_hist* hist, void* edge, const int16_t* traces, uint32_t frames,
                const hwpa_lane_model* models, uint32_t lanes, hipStream_t s, bool wrap = true);
int32_t HistAddWrap(hwpa_lane_hist* hist, void* edge, const int16_t* traces, uint32_t frames,
                    const hwpa_lane_model* models, uint32_t lanes, hipStream_t s);

// Kernel shape variants chosen per (lanes, frames) by the autotuner; every
// variant computes the same bits.  Variant 0 is the default.
enum TuneKernel { kTuneHistogram = 0, kTuneBaseline = 1, kTuneLabel = 2, kTunePulse = 3, kTuneKernels = 4 };
int32_t TuneVariants(int kernel);
int32_t TuneGet(int kernel, uint32_t lanes, uint32_t frames);
void TuneSet(int kernel, uint32_t lanes, uint32_t frames, int32_t variant);
int32_t HistAddVariant(int32_t variant, hwpa_lane_hist* hist, void* edge, const int16_t* traces, uint32_t frames,
                       const hwpa_lane_model* models, uint32_t lanes, hipStream_t s, bool wrap = true);

int32_t DmeInitModels(const hwpa_config& cfg, const hwpa_baseliner_stats* stats, hwpa_lane_model* models,
                      uint32_t lanes, hipStream_t s);
int32_t DmeEstimate(const hwpa_config& cfg, const hwpa_lane_hist* hist, const hwpa_baseliner_stats* stats,
                    hwpa_lane_model* models, const int32_t model_fi[2], const int32_t est_fi[2], uint32_t lanes,
                    int32_t* d_iters, hipStream_t s, float* diag = nullptr);

// Learned start of the EM: network input/output and fallback compaction buffers
// for up to `lanes` lanes.
struct LearnedWs {
    uint32_t lanes = 0;
    float* y = nullptr;               // [Z][11] network output
    uint8_t* fallback = nullptr;      // [Z]
    uint32_t* cnt = nullptr;          // [lanes] fallback ZMWs per lane
    uint32_t* off = nullptr;          // [lanes] exclusive prefix over lanes with 1..63 fallback ZMWs
    uint32_t* total = nullptr;        // [0] compacted fallback ZMWs, [1] all fallback ZMWs
    uint32_t* map = nullptr;          // compact slot -> lane * 64 + zmw
    hwpa_lane_hist* cHist = nullptr;  // compacted fallback inputs
    hwpa_baseliner_stats* cStats = nullptr;
    hwpa_lane_model* cModels = nullptr;
    int32_t* cIters = nullptr;
};
int32_t LearnedWsAlloc(LearnedWs& w, uint32_t lanes);
void LearnedWsFree(LearnedWs& w);
int32_t DmeEstimateLearned(const hwpa_config& cfg, hwpa_initializer* mlp, LearnedWs& w, const hwpa_lane_hist* hist,
                           const hwpa_baseliner_stats* stats, hwpa_lane_model* models, const int32_t model_fi[2],
                           const int32_t est_fi[2], uint32_t lanes, int32_t* d_iters, hipStream_t s);
int32_t MlpFeatures(const hwpa_lane_hist* hist, uint32_t lanes, uint16_t* x, hipStream_t s);
// Network outputs for every ZMW of `lanes` lane histograms (features computed in the kernel).
int32_t MlpForwardFromHist(const hwpa_initializer* m, const hwpa_lane_hist* hist, uint32_t lanes, float* y,
                           hipStream_t s);
int32_t MlpForward(const hwpa_initializer* m, const uint16_t* x, uint32_t rows, float* y, uint16_t* h1, uint16_t* h2,
                   hipStream_t s);
const float* MlpOutputMap(const hwpa_initializer* m);  // device {off[6], scale[6], w[8]}
// Logit of the learned gate's probability threshold (host): the network's confidence
// output is compared with it; a threshold of 0 lets every ZMW through.
inline float MlpGateLogit(float p) { return hw_gate_logit(p); }
int32_t MlpTargets(const hwpa_lane_hist* hist, const float* diag, uint32_t lanes, float minConf, float* t,
                   uint8_t* valid, hipStream_t s);

int32_t MetricsCreate(const hwpa_config& cfg, uint32_t lanes, hwpa_metrics** out);
void MetricsDestroy(hwpa_metrics* m);
int32_t MetricsSetTree(hwpa_metrics* m, const hwpa_activity_tree* t);
int32_t MetricsProcess(hwpa_metrics* m, const hwpa_pulse* pulses, const uint32_t* counts, uint32_t maxCalls,
                       uint32_t frames, int32_t lastFrame, const hwpa_baseliner_stats* blStats, const float* score,
                       const hwpa_stat_acc* pdStats, const hwpa_lane_model* models, hwpa_basecalling_metrics* out,
                       int32_t* produced, hipStream_t s);

struct LabelerDev;
int32_t LabelerCreate(const hwpa_config& cfg, uint32_t lanes, uint32_t max_frames, LabelerDev** out);
void LabelerDestroy(LabelerDev* l);
int32_t LabelerRun(LabelerDev* l, const int16_t* traces, uint32_t frames, const hwpa_lane_model* models,
                   int16_t* labels, int16_t* latent_out, float* score, hipStream_t s);

struct PulserDev;
int32_t PulserCreate(const hwpa_config& cfg, uint32_t lanes, PulserDev** out);
void PulserDestroy(PulserDev* p);
int32_t PulserRun(PulserDev* p, const int16_t* labels, const int16_t* latent, const int16_t* traces,
                  uint32_t frames, int32_t first_frame, const hwpa_lane_model* models, hwpa_pulse* pulses,
                  uint32_t* counts, uint32_t* overflows, hwpa_stat_acc* bl_stats, hipStream_t s);
uint32_t PulserMaxCalls(const PulserDev* p);

int32_t TracegenRun(const hwpa_tracegen_params& p, uint32_t lanes, uint32_t lane_offset, uint32_t frames,
                    uint32_t first_frame, hwpa_tracegen_state* state, int32_t init, int16_t* out,
                    hipStream_t s);

}  // namespace hwpa
