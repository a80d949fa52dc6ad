// oracle.cpp -- CPU oracle for the trace-analysis path (TEST INFRASTRUCTURE).
//
// A scalar restatement of the reference's Host* strategies.  Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg load it (as the
// checker and the CPU baseline); the product never does.
//
// Lane semantics: the reference runs every stage on 64-ZMW LaneArrays, so a
// few decisions are lane-wide (all()/any()).  Those are kept lane-wide here;
// everything else is a per-ZMW scalar loop with the reference's operation
// order.  Arithmetic follows the SSE LaneArray rules (int16 wrap, RNE float
// to int, SSE min/max NaN rule) via include/hwpa_math.h.
//
// Parity pins for each part are listed in the test that covers it.

#include "oracle.h"
#include "../include/hwpa.h"
#include "../include/hwpa_math.h"
#include "../include/hwpa_tracegen.h"

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

namespace {

constexpr int LS = HWPA_LANE_SIZE;
constexpr int NA = HWPA_NUM_ANALOGS;
constexpr int NB = HWPA_NUM_BINS;
constexpr int NS = HWPA_NUM_STATES;
// Boundaries per exact restart of the E-step density recurrence (same as the device).
constexpr int kRecurBlock = 10;

// The E-step's log / reciprocal tables (hw_logtab_entry), as the device builds them.
struct LogTabs {
    float rc[HW_LOGTAB_N], lc[HW_LOGTAB_N];
    LogTabs()
    {
        for (int k = 0; k < HW_LOGTAB_N; ++k) hw_logtab_entry(k, &rc[k], &lc[k]);
    }
};
static const LogTabs& LogTab()
{
    static const LogTabs t;
    return t;
}

int g_threads = 1;

template <typename F>
void ParLanes(int lanes, F&& fn)
{
    const int nt = std::max(1, std::min(g_threads, lanes));
    if (nt == 1) {
        for (int l = 0; l < lanes; ++l) fn(l);
        return;
    }
    std::atomic<int> next{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; ++t) {
        pool.emplace_back([&]() {
            for (int l = next.fetch_add(1); l < lanes; l = next.fetch_add(1)) fn(l);
        });
    }
    for (auto& th : pool) th.join();
}

// ---------------------------------------------------------------------------
// Configuration-derived constants
// ---------------------------------------------------------------------------

// traceAnalysis/BaselinerParams.cpp:10-72 and BaselinerParams.h:45-70
hwpa_filter_params FilterParams(int type)
{
    hwpa_filter_params p{};
    auto set = [&](std::initializer_list<int> s, std::initializer_list<int> w, float sb, float mb) {
        p.num_stages = (int)s.size();
        int i = 0;
        for (int v : s) p.strides[i++] = v;
        i = 0;
        for (int v : w) p.widths[i++] = v;
        p.sigma_bias = sb;
        p.mean_bias = mb;
    };
    switch (type) {
    case HWPA_FILTER_MULTISCALE_LARGE: set({1, 2, 8}, {11, 17, 61}, 2.64f, 0.78f); break;
    case HWPA_FILTER_MULTISCALE_MEDIUM: set({1, 2, 8}, {9, 17, 31}, 2.41f, 0.69f); break;
    case HWPA_FILTER_MULTISCALE_SMALL: set({1, 2, 8}, {7, 17, 17}, 2.15f, 0.63f); break;
    case HWPA_FILTER_TWOSCALE_LARGE: set({2, 8}, {11, 61}, 2.71f, 0.60f); break;
    case HWPA_FILTER_TWOSCALE_MEDIUM: set({2, 8}, {9, 31}, 2.44f, 0.50f); break;
    default: set({2, 8}, {7, 17}, 2.07f, 0.41f); break;
    }
    int cum = 1, minSize = 0;
    for (int i = 0; i < p.num_stages; ++i) {
        cum *= p.strides[i];
        minSize += p.widths[i] * cum;
    }
    int ret = 1;
    while (ret < minSize) ret *= 2;
    p.latent_size = ret;
    p.aggregate_stride = cum;
    return p;
}

struct DmeConsts {
    float blVarMin, blVarMax, blVarNominal, signalScaler, shotVarCoeff;
};

// traceAnalysis/CoreDMEstimator.h:67-109, CoreDMEstimator.cpp:36-87
DmeConsts MakeDmeConsts(const hwpa_config& c)
{
    DmeConsts d{};
    d.signalScaler = c.movie.photoelectron_sensitivity;
    d.shotVarCoeff = c.movie.shot_var_coeff;
    const float ss = std::max(d.signalScaler, 1.0f);
    d.blVarMin = ss * ss * (1.0f / 12.0f);
    const float t = 2.0f * d.signalScaler;
    d.blVarNominal = t * t;
    d.blVarMax = 36.0f * d.blVarNominal;
    return d;
}

// DetectionModelHost::ModelSignalCovar (traceAnalysis/DetectionModelHost.cpp:247-256)
inline float ModelSignalCovar(const DmeConsts& d, float cv2, float mean, float blVar)
{
    float r = blVar;
    r += d.shotVarCoeff * mean;
    r += cv2 * (mean * mean);
    return r;
}

// ---------------------------------------------------------------------------
// Stat accumulators (common/StatAccumulator.h, common/AutocorrAccumulator.cpp,
// dataTypes/BaselinerStatAccumulator.cpp)
// ---------------------------------------------------------------------------

struct StatZ {
    float off = 0, m0 = 0, m1 = 0, m2 = 0;
    void Add(float v)
    {
        const float x = v - off;
        m0 += 1.0f;
        m1 += x;
        m2 += x * x;
    }
    void AddMasked(float v, bool mask)
    {
        const float x = v - off;
        if (mask) {
            m0 = m0 + 1.0f;
            m1 = m1 + x;
            m2 = m2 + x * x;
        }
    }
    float Mean() const { return m1 / m0 + off; }
    float Variance() const
    {
        float var = m1 * m1 / m0;
        var = (m2 - var) / (m0 - 1.0f);
        var = hw_maxf(var, 0.0f);
        return (m0 > 1.0f) ? var : NAN;
    }
};

// AutocorrAccumulator<LaneArray<float>> with lag 4 (common/AutocorrAccumulator.{h,cpp}):
// per-ZMW moments and lag product, lane-wide buffer indices.
struct AcLane {
    StatZ s[LS];
    float m2[LS];
    float f[4][LS];
    float b[4][LS];
    int fbi = 0, bbi = 0;

    AcLane() { Reset(); }
    // AutocorrAccumulator::Reset (common/AutocorrAccumulator.h:164-171): the offset stays.
    void Reset()
    {
        for (int z = 0; z < LS; ++z) {
            s[z].m0 = s[z].m1 = s[z].m2 = 0;
            m2[z] = 0;
            for (int k = 0; k < 4; ++k) f[k][z] = b[k][z] = 0;
        }
        fbi = bbi = 0;
    }
    // AutocorrAccumulator::AddSample (common/AutocorrAccumulator.cpp:29-38)
    void AddSample(const float* v)
    {
        const bool fillFront = fbi < 4;
        const int bslot = bbi % 4;
        for (int z = 0; z < LS; ++z) {
            const float x = v[z] - s[z].off;
            if (fillFront) f[fbi][z] = x;
            m2[z] += b[bslot][z] * x;
            b[bslot][z] = x;
            s[z].Add(v[z]);
        }
        if (fillFront) ++fbi;
        bbi = (bbi + 1) % 4;
    }
    // AutocorrAccumulator::Merge (common/AutocorrAccumulator.cpp:72-102), "this" on the left.
    void Merge(const AcLane& o)
    {
        const int n1 = 4 - o.fbi;
        const int n2 = 4 - fbi;
        for (int z = 0; z < LS; ++z) {
            s[z].m0 += o.s[z].m0;
            s[z].m1 += o.s[z].m1;
            s[z].m2 += o.s[z].m2;
            m2[z] += o.m2[z];
            for (int k = 0; k < 4 - n1; ++k) {
                m2[z] += b[(bbi + k) % 4][z] * o.f[k][z];
                b[(bbi + k) % 4][z] = o.b[(o.bbi + n1 + k) % 4][z];
            }
            for (int k = 0; k < n2; ++k) f[fbi + k][z] = o.f[k][z];
        }
        fbi += n2;
        bbi = (bbi + (4 - n1)) % 4;
    }
    // AutocorrAccumulator::Autocorrelation (common/AutocorrAccumulator.cpp:42-67): the
    // clamp keeps a NaN quotient as -1 (max(NaN, -1) = -1), only nmk < 1 gives NaN.
    float Autocorrelation(int z) const
    {
        const float nmk = s[z].m0 - 4.0f;
        const float mu = s[z].Mean() - s[z].off;
        float m1x2 = 2.0f * s[z].m1;
        for (int k = 0; k < 4; ++k) m1x2 -= f[k][z] + b[k][z];
        float ac = mu * (m1x2 - nmk * mu);
        ac = (m2[z] - ac) / ((nmk - 1.0f) * s[z].Variance());
        ac = hw_minf(hw_maxf(ac, -1.0f), 1.0f);
        return (nmk < 1.0f || std::isnan(ac)) ? NAN : ac;
    }
    void ToState(hwpa_autocorr_acc* a) const
    {
        for (int z = 0; z < LS; ++z) {
            a->basic.offset[z] = s[z].off;
            a->basic.m0[z] = s[z].m0;
            a->basic.m1[z] = s[z].m1;
            a->basic.m2[z] = s[z].m2;
            a->lag_m2[z] = m2[z];
            for (int k = 0; k < 4; ++k) {
                a->fbuf[k][z] = f[k][z];
                a->bbuf[k][z] = b[k][z];
            }
            a->bidx[0][z] = (uint8_t)fbi;
            a->bidx[1][z] = (uint8_t)bbi;
        }
    }
    void FromState(const hwpa_autocorr_acc& a)
    {
        for (int z = 0; z < LS; ++z) {
            s[z] = StatZ{a.basic.offset[z], a.basic.m0[z], a.basic.m1[z], a.basic.m2[z]};
            m2[z] = a.lag_m2[z];
            for (int k = 0; k < 4; ++k) {
                f[k][z] = a.fbuf[k][z];
                b[k][z] = a.bbuf[k][z];
            }
        }
        fbi = a.bidx[0][0];
        bbi = a.bidx[1][0];
    }
};

struct LaneStats {
    AcLane ac;  // over all frames
    int16_t tmin[LS], tmax[LS];
    StatZ bl[LS];
    float rawBlSum[LS];
    StatZ bg[LS];

    LaneStats() { Reset(); }
    void Reset()
    {
        ac = AcLane{};
        for (int z = 0; z < LS; ++z) {
            tmin[z] = 32767;
            tmax[z] = -32768;
            bl[z] = StatZ{};
            rawBlSum[z] = 0;
            bg[z] = StatZ{};
        }
    }

    // BaselinerStatAccumulator::AddSample (dataTypes/BaselinerStatAccumulator.cpp:38-53)
    void AddSample(const int16_t* raw, const int16_t* bls, const bool* mask)
    {
        float v[LS];
        for (int z = 0; z < LS; ++z) v[z] = (float)bls[z];
        ac.AddSample(v);
        for (int z = 0; z < LS; ++z) {
            tmin[z] = std::min(bls[z], tmin[z]);
            tmax[z] = std::max(bls[z], tmax[z]);
            bl[z].AddMasked(v[z], mask[z]);
            rawBlSum[z] += (float)(mask[z] ? raw[z] : (int16_t)0);
        }
    }

    void AddBackground(const int16_t* v)
    {
        for (int z = 0; z < LS; ++z) bg[z].Add((float)v[z]);
    }

    void ToState(hwpa_baseliner_stats* s) const
    {
        ac.ToState(&s->full);
        for (int z = 0; z < LS; ++z) {
            s->trace_min[z] = tmin[z];
            s->trace_max[z] = tmax[z];
            s->baseline.offset[z] = bl[z].off;
            s->baseline.m0[z] = bl[z].m0;
            s->baseline.m1[z] = bl[z].m1;
            s->baseline.m2[z] = bl[z].m2;
            s->raw_baseline_sum[z] = rawBlSum[z];
            s->background.offset[z] = bg[z].off;
            s->background.m0[z] = bg[z].m0;
            s->background.m1[z] = bg[z].m1;
            s->background.m2[z] = bg[z].m2;
        }
    }

    void FromState(const hwpa_baseliner_stats* s)
    {
        ac.FromState(s->full);
        for (int z = 0; z < LS; ++z) {
            tmin[z] = s->trace_min[z];
            tmax[z] = s->trace_max[z];
            bl[z] = StatZ{s->baseline.offset[z], s->baseline.m0[z], s->baseline.m1[z], s->baseline.m2[z]};
            rawBlSum[z] = s->raw_baseline_sum[z];
            bg[z] = StatZ{s->background.offset[z], s->background.m0[z], s->background.m1[z], s->background.m2[z]};
        }
    }

    // BaselinerStatAccumulator::Merge (dataTypes/BaselinerStatAccumulator.cpp:64-75) with
    // AutocorrAccumulator::Merge and StatAccumulator::Merge (common/StatAccumulator.h:201-210).
    void Merge(const LaneStats& o)
    {
        ac.Merge(o.ac);
        for (int z = 0; z < LS; ++z) {
            tmin[z] = std::min(tmin[z], o.tmin[z]);
            tmax[z] = std::max(tmax[z], o.tmax[z]);
            bl[z].m0 += o.bl[z].m0;
            bl[z].m1 += o.bl[z].m1;
            bl[z].m2 += o.bl[z].m2;
            rawBlSum[z] += o.rawBlSum[z];
            bg[z].m0 += o.bg[z].m0;
            bg[z].m1 += o.bg[z].m1;
            bg[z].m2 += o.bg[z].m2;
        }
    }
};

// ---------------------------------------------------------------------------
// Baseliner: van Herk / Gil-Werman extremum filters
// (traceAnalysis/TraceFilters.cpp:10-57, WindowBuffer.h:12-41, BlockFilterStage.*)
// ---------------------------------------------------------------------------

struct HgwFilter {
    int W = 0, stride = 1, nf = 1;
    bool isMax = false;
    std::vector<int16_t> ring;  // [W][LS], ring[(head + j) % W] is element j
    int head = 0;
    int counter = 0;
    int16_t hold[LS];

    void Init(int w, int s, int n, bool mx)
    {
        W = w;
        stride = s;
        nf = n;
        isMax = mx;
        ring.assign((size_t)W * LS, 0);  // BlockFilterStage ctor pushes W zeros
        head = 0;
        counter = 0;  // W pushes leave the counter at W % W
        for (int z = 0; z < LS; ++z) hold[z] = 0;
    }
    int16_t Op(int16_t a, int16_t b) const { return isMax ? std::max(a, b) : std::min(a, b); }
    int16_t* R(int j) { return &ring[(size_t)((head + j) % W) * LS]; }

    // BlockFilterStage::operator() + extremumFilterHgw, in place on buf[frames][LS].
    void Run(int16_t* buf, int frames)
    {
        const int dist = frames / nf;
        int16_t s[LS];
        for (int z = 0; z < LS; ++z) s[z] = hold[z];
        int k = 0;
        for (int i = 0; i < dist; i += stride) {
            if (counter == 0) {
                for (int j = W - 1; j > 0; --j) {
                    int16_t* rj = R(j);
                    int16_t* rj1 = R(j - 1);
                    for (int z = 0; z < LS; ++z) rj1[z] = Op(rj[z], rj1[z]);
                }
                const int16_t* back = R(W - 1);
                for (int z = 0; z < LS; ++z) s[z] = back[z];
            } else {
                const int16_t* back = R(W - 1);
                for (int z = 0; z < LS; ++z) s[z] = Op(s[z], back[z]);
            }
            int16_t front[LS];
            std::memcpy(front, R(0), sizeof(front));
            int16_t* slot = &ring[(size_t)head * LS];  // push_back overwrites the front slot
            const int16_t* x = &buf[(size_t)i * LS];
            for (int z = 0; z < LS; ++z) slot[z] = x[z];
            head = (head + 1) % W;
            counter = (counter + 1) % W;
            int16_t* o = &buf[(size_t)k * LS];
            for (int z = 0; z < LS; ++z) o[z] = Op(s[z], front[z]);
            ++k;
        }
        for (int z = 0; z < LS; ++z) hold[z] = s[z];
    }
};

struct LaneBaseliner {
    std::vector<HgwFilter> lower, upper;
    float emaSum[LS], emaWeight[LS], sigmaEma[LS];
    int16_t latData[LS], latRaw[LS];
    bool latL[LS], latH1[LS], latH2[LS];
};

struct BaselinerCtx {
    hwpa_filter_params fp;
    float cMeanBias, cSigmaBias, meanEmaAlpha, sigmaEmaAlpha, jumpTol, scaler;
    int16_t pedestal;
    int lanes;
    std::vector<LaneBaseliner> lb;
};

// HostMultiScaleBaseliner::LaneBaseliner::GetSmoothedBlEstimate
// (traceAnalysis/HostMultiScaleBaseliner.cpp:178-212)
float SmoothedBlEstimate(const BaselinerCtx& c, LaneBaseliner& L, int z, int16_t lower, int16_t upper)
{
    const float minSigma = .288675135f;
    const float sigma = hw_maxf((float)hw_wrap16((int32_t)upper - (int32_t)lower) / c.cSigmaBias, minSigma);
    const float newSigmaEma = c.sigmaEmaAlpha * L.sigmaEma[z] + (1.0f - c.sigmaEmaAlpha) * sigma;
    const float blEst = 0.5f * (float)hw_wrap16((int32_t)upper + (int32_t)lower) + c.cMeanBias * newSigmaEma;
    const bool mask = (L.emaWeight[z] == 0.0f) ||
                      (blEst - L.emaSum[z] / L.emaWeight[z] < c.jumpTol * L.sigmaEma[z]);
    const float newWeight = c.meanEmaAlpha * L.emaWeight[z] + (1.0f - c.meanEmaAlpha);
    const float newSum = c.meanEmaAlpha * L.emaSum[z] + (1.0f - c.meanEmaAlpha) * blEst;
    if (mask) {
        L.emaWeight[z] = newWeight;
        L.emaSum[z] = newSum;
        L.sigmaEma[z] = newSigmaEma;
    }
    return L.emaSum[z] / L.emaWeight[z];
}

// LaneBaseliner::EstimateBaseline + AddToBaselineStats
// (traceAnalysis/HostMultiScaleBaseliner.cpp:93-176)
void EstimateBaselineLane(BaselinerCtx& c, int l, const int16_t* raw, int frames, int16_t* out,
                          hwpa_baseliner_stats* statsOut)
{
    LaneBaseliner& L = c.lb[l];
    std::vector<int16_t> lo(raw, raw + (size_t)frames * LS), up(raw, raw + (size_t)frames * LS);
    for (auto& f : L.lower) f.Run(lo.data(), frames);
    for (auto& f : L.upper) f.Run(up.data(), frames);

    LaneStats st;
    const int stride = c.fp.aggregate_stride;
    const int fd = frames / stride;
    const float thrScale = 4.5f;
    for (int i = 0; i < fd; ++i) {
        float subBl[LS];
        for (int z = 0; z < LS; ++z) {
            const int16_t lv = hw_wrap16((int32_t)lo[(size_t)i * LS + z] - c.pedestal);
            const int16_t uv = hw_wrap16((int32_t)up[(size_t)i * LS + z] - c.pedestal);
            const float blEst = SmoothedBlEstimate(c, L, z, lv, uv);
            subBl[z] = blEst * c.scaler;
        }
        for (int j = 0; j < stride; ++j) {
            const size_t t = (size_t)(i * stride + j);
            int16_t rawRounded[LS], bls[LS];
            bool mask[LS];
            for (int z = 0; z < LS; ++z) {
                const float rawScaled = (float)hw_wrap16((int32_t)raw[t * LS + z] - c.pedestal) * c.scaler;
                bls[z] = hw_round_i16(rawScaled - subBl[z]);
                out[t * LS + z] = bls[z];
                rawRounded[z] = hw_round_i16(rawScaled);
                const float thrLow = thrScale * L.sigmaEma[z];
                const float thrHigh = thrScale * L.sigmaEma[z];
                const bool maskHp1 = (float)bls[z] < thrHigh * c.scaler;
                mask[z] = L.latH1[z] && L.latL[z] && maskHp1;
                L.latL[z] = (float)bls[z] < thrLow * c.scaler;
                L.latH2[z] = L.latH1[z];
                L.latH1[z] = maskHp1;
            }
            st.AddSample(L.latRaw, L.latData, mask);
            for (int z = 0; z < LS; ++z) {
                L.latRaw[z] = rawRounded[z];
                L.latData[z] = bls[z];
            }
        }
        int16_t bgv[LS];
        for (int z = 0; z < LS; ++z) bgv[z] = hw_round_i16(subBl[z]);
        st.AddBackground(bgv);
    }
    st.ToState(statsOut);
}

// ---------------------------------------------------------------------------
// Trace histogram (traceAnalysis/TraceHistogramAccumHost.cpp:92-169,
// dataTypes/UHistogramSimd.*, traceAnalysis/EdgeFrameClassifier.h)
// ---------------------------------------------------------------------------

struct LaneHistState {
    float binSize[LS];
    float binStart[NB + 1][LS];
    uint16_t lowOut[LS], highOut[LS];
    uint16_t count[NB + 1][LS];
    // EdgeFrameClassifier: never reset by histogram resets
    int16_t prevFrame[LS];
    bool hist[2][LS];
    int hi0 = 0, hi1 = -1;
};

struct HistCtx {
    float binSizeCoeff, binSizeMin, fallbackSigma;
    float minFrameCount;
    int lanes;
    std::vector<LaneHistState> h;
};

// TraceHistogramAccumHost::ResetImpl(BaselinerMetrics) + UHistogramSimd ctor
int HistResetLane(HistCtx& c, int l, const hwpa_baseliner_stats& s)
{
    LaneHistState& H = c.h[l];
    int err = 0;
    for (int z = 0; z < LS; ++z) {
        StatZ st{s.baseline.offset[z], s.baseline.m0[z], s.baseline.m1[z], s.baseline.m2[z]};
        const bool suff = st.m0 >= c.minFrameCount;
        const float mean = suff ? st.Mean() : 0.0f;
        const float sigma = suff ? hw_sqrtf(st.Variance()) : c.fallbackSigma;
        const float binSize = hw_maxf(c.binSizeCoeff * sigma, c.binSizeMin);
        const float lo = mean - 4.0f * sigma;
        const float up = lo + (float)NB * binSize;
        if (!(lo < up)) err = 1;
        const float bs = (up - lo) / (float)NB;
        if (!(bs > 0.0f)) err = 1;
        H.binSize[z] = bs;
        H.binStart[0][z] = lo;
        for (int i = 1; i <= NB; ++i) H.binStart[i][z] = bs * (float)i + lo;
        H.binStart[NB][z] = up;
        H.lowOut[z] = H.highOut[z] = 0;
        for (int i = 0; i <= NB; ++i) H.count[i][z] = 0;
    }
    return err;
}

void HistAddLane(HistCtx& c, int l, const int16_t* tr, int frames, const hwpa_lane_model& m)
{
    LaneHistState& H = c.h[l];
    int16_t thr[LS];
    for (int z = 0; z < LS; ++z) {
        const float var = hw_h2f(m.baseline.vars[z]);
        const float mean = hw_h2f(m.baseline.means[z]);
        thr[z] = hw_wrap16(hw_round_i32(2.0f * hw_sqrtf(var) + mean));
    }
    for (int t = 0; t < frames; ++t) {
        const int16_t* f = tr + (size_t)t * LS;
        const bool full = (H.hi1 >= 0) && (1 + (2 + H.hi1 - H.hi0) % 2) == 2;
        bool isBl[LS], edge[LS];
        for (int z = 0; z < LS; ++z) {
            isBl[z] = f[z] < thr[z];
            edge[z] = full ? (isBl[z] ^ H.hist[H.hi0][z]) : true;
        }
        // CircularArray<LaneMask,2>::push_back (common/CircularArray.h)
        if (full) H.hi0 = (H.hi0 + 1) % 2;
        H.hi1 = (H.hi1 + 1) % 2;
        for (int z = 0; z < LS; ++z) H.hist[H.hi1][z] = isBl[z];
        for (int z = 0; z < LS; ++z) {
            const int16_t cand = H.prevFrame[z];
            H.prevFrame[z] = f[z];
            if (edge[z]) continue;
            const float y = ((float)cand - H.binStart[0][z]) / H.binSize[z];
            const int bin = hw_floor_i32(y);
            if (bin < 0) H.lowOut[z] = (uint16_t)(H.lowOut[z] + 1);
            else if (bin >= NB) H.highOut[z] = (uint16_t)(H.highOut[z] + 1);
            else H.count[bin][z] = (uint16_t)(H.count[bin][z] + 1);
        }
    }
}

// Checker of hw_shortbin against the divide of AddBlock (see oracle.h).
static void ShortbinSweep(const float* lo, const float* bs, int32_t n, uint64_t out[3])
{
    uint64_t inRange = 0, decided = 0, wrong = 0;
    for (int32_t c = 0; c < n; ++c) {
        const hw_shortbin_z zb = hw_shortbin_make(lo[c], bs[c]);
        for (int32_t x = -32768; x <= 32767; ++x) {
            const int exact = hw_floor_i32(((float)x - lo[c]) / bs[c]);
            int32_t b;
            const bool ok = hw_shortbin(x, zb, NB, &b);
            const bool in = exact >= 0 && exact < NB;
            inRange += in ? 1u : 0u;
            decided += (in && ok) ? 1u : 0u;
            wrong += (ok && (!in || b != exact)) ? 1u : 0u;
        }
    }
    out[0] = inRange;
    out[1] = decided;
    out[2] = wrong;
}

void HistGetLane(const HistCtx& c, int l, hwpa_lane_hist& o)
{
    const LaneHistState& H = c.h[l];
    for (int z = 0; z < LS; ++z) {
        o.low_bound[z] = H.binStart[0][z];
        o.bin_size[z] = H.binSize[z];
        o.outlier_low[z] = H.lowOut[z];
        o.outlier_high[z] = H.highOut[z];
        for (int b = 0; b < NB; ++b) o.bin_count[b][z] = H.count[b][z];
    }
}

// ---------------------------------------------------------------------------
// Detection model (traceAnalysis/DetectionModelHost.*, DmeEmHost.cpp)
// ---------------------------------------------------------------------------

struct ModeZ {
    float w, mean, var;
};
struct ModelZ {
    ModeZ bl;
    ModeZ an[NA];
    float conf;
};

ModelZ ModelFromHalf(const hwpa_lane_model& m, int z)
{
    ModelZ r{};
    r.bl = {hw_h2f(m.baseline.weights[z]), hw_h2f(m.baseline.means[z]), hw_h2f(m.baseline.vars[z])};
    for (int a = 0; a < NA; ++a)
        r.an[a] = {hw_h2f(m.analogs[a].weights[z]), hw_h2f(m.analogs[a].means[z]), hw_h2f(m.analogs[a].vars[z])};
    r.conf = hw_h2f(m.confidence[z]);
    return r;
}

// SignalModeHost::ExportTo / DetectionModelHost::ExportTo (DetectionModelHost.cpp:225-245,270-278)
void ModelToHalf(const ModelZ& r, hwpa_lane_model& m, int z)
{
    m.baseline.weights[z] = hw_f2h(r.bl.w);
    m.baseline.means[z] = hw_f2h(r.bl.mean);
    m.baseline.vars[z] = hw_f2h(hw_minf(r.bl.var, 65000.0f));
    for (int a = 0; a < NA; ++a) {
        m.analogs[a].weights[z] = hw_f2h(r.an[a].w);
        m.analogs[a].means[z] = hw_f2h(r.an[a].mean);
        m.analogs[a].vars[z] = hw_f2h(hw_minf(r.an[a].var, 65000.0f));
    }
    m.confidence[z] = hw_f2h(r.conf);
}

struct DmeCfg {
    DmeConsts k;
    float refSnr;
    float relAmp[NA], cv[NA];
    float mixThresh0, mixThresh1, scaleSnrConfTol, pulseAmpReg, snrDrop, snrThresh0, snrThresh1, successThresh;
    float hl0, hl1;
    unsigned iterLimit;
    bool iterToLimit;
    unsigned updateMethod;
    float log2pi;
};

DmeCfg MakeDmeCfg(const hwpa_config& c)
{
    DmeCfg d{};
    d.k = MakeDmeConsts(c);
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
    d.iterLimit = c.em_iteration_limit;
    d.iterToLimit = c.iterate_to_limit != 0;
    d.updateMethod = c.model_update_method;
    d.log2pi = std::log(2.0f * 3.1415926536f);   // libm, scalar (DmeEmHost.cpp:380)
    return d;
}

// The two oracle forms of the DME differ in their transcendentals and in the EM
// iteration.  ReferenceForm evaluates every function as the reference's x86
// host build does: LaneArray exp / log are sse_mathfun's Cephes routines
// (hw_cephes_expf / hw_cephes_logf), LaneArray exp2 / log2 / pow call libm per
// element (common/simd/xcompile.h:180-197, m512f.h:34-41), scalar std::log is
// libm; its EM is the log-sum-exp iteration of DmeEmHost.cpp:374-565 with a
// division per bin and no fused steps.  MirrorForm is the GPU kernels'
// arithmetic (include/hwpa_math.h routines and the restated E-step of
// EmIterationsMirror), bit for bit.
struct ReferenceForm {
    static constexpr bool kReference = true;
    static float Exp(float x) { return hw_cephes_expf(x); }
    static float Log(float x) { return hw_cephes_logf(x); }
    static float Exp2(float x) { return std::exp2(x); }
    static float Log2(float x) { return std::log2(x); }
    static float Pow(float x, float y) { return std::pow(x, y); }
    static float ScalarLog(float x) { return std::log(x); }
};
// ReferenceForm with libm's correctly rounded-to-within-an-ulp expf / logf in place
// of the Cephes routines: a restatement that differs from ReferenceForm only in the
// last bits of exp and log, the yardstick for how far any two faithful
// implementations of DmeEmHost can sit apart.
struct ReferenceLibmForm : ReferenceForm {
    static float Exp(float x) { return std::exp(x); }
    static float Log(float x) { return std::log(x); }
};
struct MirrorForm {
    static constexpr bool kReference = false;
    static float Exp(float x) { return hw_cephes_expf(x); }
    static float Log(float x) { return hw_cephes_logf(x); }
    static float Exp2(float x) { return hw_dexp2f(x); }
    static float Log2(float x) { return hw_dlog2f(x); }
    static float Pow(float x, float y) { return hw_dpowf(x, y); }
    static float ScalarLog(float x) { return std::log(x); }
};

// DetectionModelHost::Update0/1/2 (DetectionModelHost.cpp:69-165)
template <class M>
void UpdateFraction(const DmeCfg& d, ModelZ& t, const ModelZ& o, float f)
{
    const float a = 1.0f - f;
    const float b = f;
    if (d.updateMethod == 0) {
        t.bl.mean = a * t.bl.mean + b * o.bl.mean;
        t.bl.var = a * t.bl.var + b * o.bl.var;
        for (int i = 0; i < NA; ++i) {
            t.an[i].mean = a * t.an[i].mean + b * o.an[i].mean;
            t.an[i].var = a * t.an[i].var + b * o.an[i].var;
        }
        return;
    }
    const float prevBlVar = t.bl.var;
    const float bw = a * t.bl.w + b * o.bl.w;
    t.bl.w = bw;
    t.bl.mean = a * t.bl.mean + b * o.bl.mean;
    const float bv = M::Pow(t.bl.var, a) * M::Pow(o.bl.var, b);
    t.bl.var = bv;
    const float aw = 0.25f * (1.0f - bw);
    for (int i = 0; i < NA; ++i) {
        t.an[i].w = aw;
        float cv2;
        if (d.updateMethod == 2) {
            auto xsn = [&](float m, float v, float blv) {
                float r = v - blv;
                r -= m * d.k.shotVarCoeff;
                r /= (m * m);
                return r;
            };
            const float tx = xsn(t.an[i].mean, t.an[i].var, prevBlVar);
            const float ox = xsn(o.an[i].mean, o.an[i].var, o.bl.var);
            cv2 = a * tx + b * ox;
        } else {
            cv2 = d.cv[i] * d.cv[i];
        }
        const float am = M::Pow(t.an[i].mean, a) * M::Pow(o.an[i].mean, b);
        t.an[i].mean = am;
        t.an[i].var = ModelSignalCovar(d.k, cv2, am, bv);
    }
}

// DetectionModelHost::Update(const DetectionModelHost&) (DetectionModelHost.cpp:183-205)
template <class M>
void UpdateConf(const DmeCfg& d, ModelZ& t, const ModelZ& o)
{
    const float confSum = t.conf + o.conf;
    const float fraction = (confSum > 0.0f) ? o.conf / confSum : 0.0f;
    UpdateFraction<M>(d, t, o, fraction);
    t.conf = confSum;
}

// DmeEmHost::InitLaneDetModel (DmeEmHost.cpp:871-925)
void InitLaneModel(const DmeCfg& d, const hwpa_baseliner_stats& s, hwpa_lane_model& m)
{
    for (int z = 0; z < LS; ++z) {
        StatZ bst{s.baseline.offset[z], s.baseline.m0[z], s.baseline.m1[z], s.baseline.m2[z]};
        const float blWeight = s.baseline.m0[z] / s.full.basic.m0[z];
        float blMean = bst.Mean();
        float blVar = bst.Variance();
        if (hw_isnan(blMean)) blMean = 0.0f;
        if (hw_isnan(blVar)) blVar = 100.0f;
        blVar = hw_clampf(blVar, d.k.blVarMin, d.k.blVarMax);
        m.confidence[z] = hw_f2h(0.1f);
        m.baseline.weights[z] = hw_f2h(blWeight);
        m.baseline.means[z] = hw_f2h(blMean);
        m.baseline.vars[z] = hw_f2h(blVar);
        const float amw = (1.0f - blWeight) / (float)NA;
        const float refSignal = d.refSnr * hw_sqrtf(blVar);
        for (int a = 0; a < NA; ++a) {
            m.analogs[a].weights[z] = hw_f2h(amw);
            const float aMean = hw_maxf(0.0f, blMean + d.relAmp[a] * refSignal);
            m.analogs[a].means[z] = hw_f2h(aMean);
            m.analogs[a].vars[z] = hw_f2h(ModelSignalCovar(d.k, d.cv[a] * d.cv[a], aMean, blVar));
        }
    }
}

// Histogram view for the EM (UHistogramSimd(const LaneHistogram&), UHistogramSimd.cpp:100-118)
struct HistZ {
    float binSize;
    float start[NB + 1];
    uint16_t cnt[NB + 1];  // cnt[NB] is UHistogramSimd's always-empty overflow bin
    uint16_t lowOut, highOut;
    uint16_t InRange() const
    {
        uint16_t s = 0;
        for (int i = 0; i < NB; ++i) s = (uint16_t)(s + cnt[i]);
        return s;
    }
    uint16_t Total() const { return (uint16_t)(InRange() + (uint16_t)(lowOut + highOut)); }

    // UHistogramSimd::CumulativeCount (UHistogramSimd.cpp:199-238)
    float CumulativeCount(float x) const
    {
        const bool mNan = hw_isnan(x);
        const bool mLow = x < start[0];
        const bool mHigh = x >= start[NB];
        const bool mIn = !(mNan || mLow || mHigh);
        int xbin = hw_floor_i32((x - start[0]) / binSize);
        if (!mIn) xbin = 0;
        const float xrem = x - start[xbin];
        const float xbinCount = (float)cnt[xbin];
        float cc = (float)lowOut;
        uint16_t below = 0;
        for (int i = 0; i < xbin; ++i) below = (uint16_t)(below + cnt[i]);
        cc += (float)below;
        cc += xbinCount * xrem / binSize;
        if (mLow) cc = (float)lowOut;
        if (mHigh) cc = (float)(uint16_t)(lowOut + InRange());
        if (mNan) cc = NAN;
        return cc;
    }

    // UHistogramSimd::Fractile (UHistogramSimd.cpp:140-197)
    float Fractile(float frac) const
    {
        const float nfz = frac * (float)Total();
        uint16_t n = lowOut;
        if (n > 0 && (float)n >= nfz) return -INFINITY;
        int i = 0;
        while ((n == 0 || (float)n < nfz) && i < NB) n = (uint16_t)(n + cnt[i++]);
        if ((float)n < nfz) return INFINITY;
        i -= 1;
        const float x0 = start[i];
        const uint16_t niz = cnt[i];
        const int m = (int)n - (int)niz;
        return x0 + binSize * (nfz - (float)m) / (float)(niz + 1);
    }
};

// Eigen 3.3 reduction orders used by DmeEmHost (fixed-size 5: tree; dynamic: linear).
inline float Sum5Tree(const float* v) { return (v[0] + v[1]) + (v[2] + (v[3] + v[4])); }

struct EmZ {
    ModelZ det;      // detModel (running)
    ModelZ work;     // workModel
    float scaleFactor;
    float rho[5], mu[5], var[5];
    float rhoEst[5], muEst[5], varEst[5];
    float s, sExpect, sExpectWeight;
    float logLike, logLikePrev;
    float binProbMin;
    int status;
    bool converged;
    int iterCount;      // first converged iteration
    int captureIter;    // last iteration whose parameters became the estimate (-1: none)
    float mlLogLike, mlDelta;
    uint16_t numFrames;
    float njSum;
    float varQuant;
    float piv, q0, q1, q2;  // baseline moments about the pivot, E-step -> M-step
};

// DmeEmHost::PrelimScaleFactor (DmeEmHost.cpp:618-654)
float PrelimScaleFactor(const DmeCfg& d, const ModelZ& m, const HistZ& h)
{
    const float bgSigma = hw_sqrtf(m.bl.var);
    const float thresh = 2.5f * bgSigma + m.bl.mean;
    const float bgCount = h.CumulativeCount(thresh);
    const float totalCount = (float)(uint16_t)(h.lowOut + h.InRange());
    const float fractile = 0.5f * (bgCount + totalCount) / totalCount;
    float avg = 0.0f;
    for (int a = 0; a < NA; ++a) avg += m.an[a].mean;
    avg /= (float)NA;
    const float w = hw_satlin(0.0f, d.scaleSnrConfTol, m.conf);
    float sf = h.Fractile(fractile) / avg;
    sf = (1.0f - w) * sf + w;
    return hw_minf(hw_maxf(sf, 0.1f), 10.0f);
}

// DmeEmHost::ComputeConfidence (DmeEmHost.cpp:717-803)
template <class M>
void ComputeConfidence(const DmeCfg& d, bool converged, const ModelZ& ref, const ModelZ& est, float* cf)
{
    cf[0] = converged ? 1.0f : 0.0f;
    cf[1] = hw_satlin(0.05f, 0.15f, est.bl.w);
    float x = (est.bl.mean * est.bl.mean) / est.bl.var;
    x = M::Exp(-x / (2.0f * (1.0f * 1.0f)));
    cf[2] = x;
    x = M::Log2(est.bl.var / ref.bl.var);
    const float tol = 1.5f / (0.5f + ref.conf);
    x = M::Exp(-x * x / (2.0f * (tol * tol)));
    cf[3] = x;
    x = 1.0f;
    if (d.mixThresh1 > 0.0f) {
        x *= hw_satlin(d.mixThresh0, d.mixThresh1, est.an[0].w);
        x *= hw_satlin(d.mixThresh0, d.mixThresh1, est.an[NA - 1].w);
    }
    cf[4] = x;
    const float bgSigma = hw_sqrtf(est.bl.var);
    x = est.an[NA - 1].mean;
    x /= bgSigma;
    cf[5] = hw_satlin(d.snrThresh0, d.snrThresh1, x);
    if (d.snrDrop < 0.0f) {
        cf[6] = 1.0f;
    } else {
        const float snrEst = est.an[0].mean / bgSigma;
        const float refBgSigma = hw_sqrtf(ref.bl.var);
        const float refSnr0 = ref.an[0].mean / refBgSigma;
        float refSnr1 = ref.an[1].mean / refBgSigma;
        refSnr1 *= d.snrDrop;
        refSnr1 *= hw_minf(ref.conf, 1.0f);
        cf[6] = hw_satlin(refSnr1, hw_sqrtf(refSnr0 * refSnr1), snrEst);
    }
    cf[7] = 1.0f;  // G-test disabled (GTestStatFactor < 0): pValue stays 1
}

// EvolveModel (DmeEmHost.cpp:806-834) + EvolveConfidence (DetectionModelHost.h:190-202):
// the confidence multiplier 2^(-|tEst - tModel| / hl) between interval centres.
template <class M>
float EvolveFactor(const DmeCfg& d, const int32_t* modelFi, const int32_t* estFi)
{
    const int32_t tModel = modelFi[0] + (int32_t)((uint32_t)(modelFi[1] - modelFi[0]) / 2u);
    const int32_t tEst = estFi[0] + (int32_t)((uint32_t)(estFi[1] - estFi[0]) / 2u);
    const float thl = 0.5f * (float)(tModel + tEst);
    const float hl = d.hl0 + (d.hl1 - d.hl0) * hw_satlin(56160.0f, 168480.0f, thl);
    const float tDiff = (float)(tEst - tModel);
    return M::Exp2(-hw_fabsf(tDiff) / hl);
}

// UHistogramSimd(const LaneHistogram&) for one ZMW (UHistogramSimd.cpp:100-118).
void MakeHistZ(const hwpa_lane_hist& lh, int z, HistZ& h)
{
    h.binSize = lh.bin_size[z];
    for (int b = 0; b <= NB; ++b) h.start[b] = lh.low_bound[z] + (float)b * h.binSize;
    for (int b = 0; b < NB; ++b) h.cnt[b] = lh.bin_count[b][z];
    h.cnt[NB] = 0;
    h.lowOut = lh.outlier_low[z];
    h.highOut = lh.outlier_high[z];
}

// The EM iterations as the GPU kernel computes them (hw_amd/csrc/dme.hip
// DmeEstimateKernel): one pass over the bin boundaries per iteration with the
// component densities carried directly (no log-sum-exp), per-boundary sums in
// fused multiply-adds, the density recurrence restarted every kRecurBlock
// boundaries, the table log and reciprocal of the bin probability, and the
// baseline second moment from moments about the iteration's starting mean.
// Returns the number of iterations run.
unsigned EmIterationsMirror(const DmeCfg& d, std::vector<EmZ>& ez, const std::vector<HistZ>& hz, const float* rpa,
                            float rpaMin, unsigned iterLimit, float convTolCoeff, const bool* gated = nullptr)
{
    unsigned it = 0;
    std::vector<float> sci((size_t)LS * 5), smom1((size_t)LS * 5), svarinv((size_t)LS * 5);
    for (; it < iterLimit; ++it) {
        // E-step for every ZMW of the lane (DmeEmHost.cpp:374-483)
        bool allDone = true;
        for (int z = 0; z < LS; ++z) {
            EmZ& e = ez[z];
            const HistZ& h = hz[z];
            float varinv[5], pref[5], nhv[5];
            for (int k = 0; k < 5; ++k) {
                varinv[k] = 1.0f / e.var[k];
                pref[k] = (hw_cephes_logf(e.rho[k]) - 0.5f * hw_cephes_logf(e.var[k])) - 0.5f * d.log2pi;
                nhv[k] = -0.5f * varinv[k];
            }
            // Every use of a responsibility is boundaryProb * tau_i = exp(t_i)
            // (DmeEmHost.cpp:431-432,547-548): a boundary carries the component
            // densities dd[i] = exp(t_i), boundaryProb is their sum.  The bin sums
            // sum_b (dd_i(b) + dd_i(b+1)) n_b / binProb_b of DmeEmHost.cpp:425-445 are
            // regrouped per boundary, sum_c dd_i(c) (f_{c-1} + f_c) with
            // f_b = n_b / binProb_b, each term one fused multiply-add.  The
            // baseline second moment (DmeEmHost.cpp:536-552) is accumulated in
            // the same pass about the pivot piv = this iteration's baseline mean
            // and shifted to the new mean in the M-step.
            const float piv = e.mu[0];
            // Densities along the boundaries: dd_i(c+1) = dd_i(c) R_i(c),
            // R_i(c+1) = R_i(c) Q_i with Q_i = exp(-bs^2 / var_i) (t_i is quadratic in
            // c), recomputed from exp every kRecurBlock boundaries; a density below
            // FLT_MIN while rising, or a ratio above e^80, takes exp at every boundary
            // of the block.  The same operations as the device E-step.
            float cur[5], rr[5], qq[5];
            for (int k = 0; k < 5; ++k) qq[k] = hw_expf((2.0f * nhv[k]) * (h.binSize * h.binSize));
            bool ex[5] = {false, false, false, false, false};
            auto tAt = [&](float x, int i) {
                const float y = x - e.mu[i];
                return hw_fmaf(y * y, nhv[i], pref[i]);
            };
            float nbs[5];
            for (int k = 0; k < 5; ++k) nbs[k] = nhv[k] * h.binSize;
            auto restart = [&](int c) {
                const float x0 = h.start[c];
                for (int i = 0; i < 5; ++i) {
                    const float y0 = x0 - e.mu[i];
                    const float t0 = hw_fmaf(y0 * y0, nhv[i], pref[i]);
                    cur[i] = hw_expf(t0);
                    const float dl = nbs[i] * hw_fmaf(2.0f, y0, h.binSize);   // t(x0 + bs) - t(x0)
                    rr[i] = hw_expf(dl);
                    ex[i] = !((cur[i] >= FLT_MIN || dl <= 0.0f) && dl <= 80.0f);
                }
            };
            auto advance = [&](int c) {
                for (int i = 0; i < 5; ++i) {
                    cur[i] = cur[i] * rr[i];
                    rr[i] = rr[i] * qq[i];
                    if (ex[i]) cur[i] = hw_expf(tAt(h.start[c], i));
                }
            };
            float probSum = 0.0f, wps = 0.0f, q1 = 0.0f, q2 = 0.0f;
            float ci[5] = {0, 0, 0, 0, 0}, mom1[5] = {0, 0, 0, 0, 0};
            const float halfBin = 0.5f * h.binSize;
            auto gather = [&](const float* dc, float x, float F) {
                const float G = x * F;
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
            float prevD[5], bpPrev = Sum5Tree(cur);
            for (int m = 0; m < 5; ++m) prevD[m] = cur[m];
            float fPrev = 0.0f;
            static_assert(NB % kRecurBlock == 0, "whole blocks of boundaries");
            for (int b = 0; b < NB; ++b) {
                if ((b + 1) % kRecurBlock == 0) restart(b + 1);
                else advance(b + 1);
                const float bpNext = Sum5Tree(cur);
                const float binp = hw_maxf(halfBin * (bpPrev + bpNext), e.binProbMin);
                probSum += binp;
                const float nj = (float)h.cnt[b];
                float lg, rcp;   // binp >= binProbMin >= FLT_MIN and finite
                hw_log_rcp_tab(binp, LogTab().rc, LogTab().lc, &lg, &rcp);
                wps = hw_fmaf(nj, lg, wps);
                const float f = nj * rcp;
                gather(prevD, h.start[b], fPrev + f);
                bpPrev = bpNext;
                for (int m = 0; m < 5; ++m) prevD[m] = cur[m];
                fPrev = f;
            }
            gather(prevD, h.start[NB], fPrev);
            e.q0 = ci[0];
            e.q1 = q1;
            e.q2 = q2;
            e.piv = piv;
            for (int m = 0; m < 5; ++m) {
                ci[m] *= halfBin;
                mom1[m] *= halfBin;
            }
            e.logLikePrev = e.logLike;
            e.logLike = wps - e.njSum * hw_cephes_logf(probSum);
            const float ds = e.s - e.sExpect;
            e.logLike -= 0.5f * (ds * ds) * e.sExpectWeight;
            const float delta = e.logLike - e.logLikePrev;
            const float convTol = convTolCoeff * hw_fabsf(e.logLike);
            bool conv = (delta >= 0.0f) && (delta < convTol);
            conv = conv && (e.logLike >= e.mlLogLike);
            // MaxLikelihoodDiagnostics::Converged (MaxLikelihoodDiagnostics.h:90-101)
            if (conv && !e.converged) e.iterCount = (int)it;
            if (conv) {
                e.mlLogLike = e.logLike;
                e.mlDelta = delta;
                e.captureIter = (int)it;
                for (int k = 0; k < 5; ++k) {
                    e.rhoEst[k] = e.rho[k];
                    e.muEst[k] = e.mu[k];
                    e.varEst[k] = e.var[k];
                }
            }
            e.converged = e.converged || conv;
            if (!(e.converged || e.status != HWPA_ZS_OK || (gated && gated[z]))) allDone = false;
            for (int m = 0; m < 5; ++m) {
                sci[(size_t)z * 5 + m] = ci[m];
                smom1[(size_t)z * 5 + m] = mom1[m];
                svarinv[(size_t)z * 5 + m] = varinv[m];
            }
        }
        // Lane-wide early exit (DmeEmHost.cpp:485)
        if (!d.iterToLimit && allDone) break;

        // M-step (DmeEmHost.cpp:494-564)
        for (int z = 0; z < LS; ++z) {
            EmZ& E = ez[z];
            const HistZ& H = hz[z];
            float c[5], mm[5], vi[5];
            for (int m = 0; m < 5; ++m) {
                c[m] = hw_maxf(sci[(size_t)z * 5 + m], 0.1f);
                mm[m] = smom1[(size_t)z * 5 + m];
                vi[m] = svarinv[(size_t)z * 5 + m];
            }
            for (int m = 0; m < 5; ++m) E.rho[m] = c[m] / E.njSum;
            const float rs = Sum5Tree(E.rho);
            for (int m = 0; m < 5; ++m) E.rho[m] = E.rho[m] / rs;
            E.mu[0] = mm[0] / c[0];
            float numer = 0.0f, denom = 0.0f;
            for (int a = 0; a < NA; ++a) {
                const int i = a + 1;
                numer += rpa[a] * vi[i] * mm[i];
                denom += (rpa[a] * rpa[a]) * vi[i] * c[i];
            }
            numer += E.sExpect * E.sExpectWeight;
            denom += E.sExpectWeight;
            E.s = numer / denom;
            const float minPulseMean = hw_maxf(1.0f, d.k.signalScaler);
            const float sMin = hw_maxf(0.5f * hw_sqrtf(E.var[0]), minPulseMean) / rpaMin;
            if (E.s < sMin) {
                E.s = sMin;
                E.status |= HWPA_ZS_VLOW_SIGNAL;
            }
            for (int a = 0; a < NA; ++a) E.mu[a + 1] = E.s * rpa[a];
            // Baseline second moment about the new mean, from the moments about the pivot.
            const float dl = E.mu[0] - E.piv;
            float mom2 = (E.q2 - (2.0f * dl) * E.q1) + (dl * dl) * E.q0;
            mom2 *= 0.5f * H.binSize;
            E.var[0] = mom2 / c[0] + E.varQuant;
            E.var[0] = hw_clampf(E.var[0], d.k.blVarMin, d.k.blVarMax);
            for (int a = 0; a < NA; ++a)
                E.var[a + 1] = ModelSignalCovar(d.k, d.cv[a] * d.cv[a], E.mu[a + 1], E.var[0]);
        }
    }
    return it;
}

// The EM iterations of DmeEmHost::EstimateLaneDetModel as the reference writes
// them (DmeEmHost.cpp:374-565), per ZMW with lane-wide exit: the E-step takes
// the log density tau_i of every mode at every bin boundary, the log-sum-exp
// llb over the modes (Eigen's fixed-size-5 reductions, (a0+a1)+(a2+(a3+a4)),
// for the max and the sum), tau_i := exp(tau_i - llb) and boundaryProb =
// exp(llb); each bin then divides its count by its trapezoid probability and
// gathers the responsibilities.  The M-step's baseline variance takes a second
// pass over the bins with the new baseline mean.  Transcendentals of M (the
// reference's Cephes exp / log), true divisions, no fused steps.
template <class M>
unsigned EmIterationsRef(const DmeCfg& d, std::vector<EmZ>& ez, const std::vector<HistZ>& hz, const float* rpa,
                         float rpaMin)
{
    constexpr int NE = NB + 1;
    static thread_local std::vector<float> tauBuf, bpBuf, binProbBuf;
    tauBuf.resize((size_t)LS * 5 * NE);
    bpBuf.resize((size_t)LS * NE);
    binProbBuf.resize((size_t)LS * NB);
    // static const float log_2pi = std::log(2.0f * pi_f) (DmeEmHost.cpp:380; libm, scalar)
    const float log2pi = M::ScalarLog(2.0f * 3.1415926536f);
    auto max5 = [](const float* v) { return hw_maxf(hw_maxf(v[0], v[1]), hw_maxf(v[2], hw_maxf(v[3], v[4]))); };
    float ci[LS][5], mom1[LS][5], varinv[LS][5];
    unsigned it = 0;
    for (; it < d.iterLimit; ++it) {
        bool allDone = true;
        for (int z = 0; z < LS; ++z) {
            EmZ& e = ez[z];
            const HistZ& h = hz[z];
            float* tau = &tauBuf[(size_t)z * 5 * NE];      // tau[i * NE + b]
            float* bp = &bpBuf[(size_t)z * NE];
            float* binProb = &binProbBuf[(size_t)z * NB];
            float pref[5];
            for (int i = 0; i < 5; ++i) {
                varinv[z][i] = 1.0f / e.var[i];
                pref[i] = (M::Log(e.rho[i]) - 0.5f * M::Log(e.var[i])) - 0.5f * log2pi;
            }
            float probSum = 0.0f, weightedProbSum = 0.0f;
            for (int i = 0; i < 5; ++i) ci[z][i] = mom1[z][i] = 0.0f;
            auto cornerComp = [&](int b) {
                const float x = h.start[b];
                float t[5];
                for (int i = 0; i < 5; ++i) {
                    const float y = x - e.mu[i];
                    t[i] = pref[i] - 0.5f * y * y * varinv[z][i];
                }
                const float tauMax = max5(t);
                float ex[5];
                for (int i = 0; i < 5; ++i) ex[i] = M::Exp(t[i] - tauMax);
                const float llb = tauMax + M::Log(Sum5Tree(ex));
                for (int i = 0; i < 5; ++i) tau[i * NE + b] = M::Exp(t[i] - llb);
                bp[b] = M::Exp(llb);
            };
            auto centerComp = [&](int b) {
                binProb[b] = hw_maxf(0.5f * h.binSize * (bp[b] + bp[b + 1]), e.binProbMin);
                probSum += binProb[b];
                const float nj = (float)h.cnt[b];
                weightedProbSum += nj * M::Log(binProb[b]);
                const float factor = nj / binProb[b];
                const float x0 = h.start[b];
                const float x1 = h.start[b + 1];
                for (int m = 0; m < 5; ++m) {
                    const float tmp1 = bp[b] * tau[m * NE + b];
                    const float tmp2 = bp[b + 1] * tau[m * NE + b + 1];
                    ci[z][m] += (tmp1 + tmp2) * factor;
                    mom1[z][m] += (tmp1 * x0 + tmp2 * x1) * factor;
                }
            };
            cornerComp(0);
            for (int b = 0; b < NB; ++b) {
                cornerComp(b + 1);
                centerComp(b);
            }
            const float hb = 0.5f * h.binSize;
            for (int m = 0; m < 5; ++m) {
                ci[z][m] *= hb;
                mom1[z][m] *= hb;
            }
            e.logLikePrev = e.logLike;
            e.logLike = weightedProbSum - e.njSum * M::Log(probSum);
            const float ds = e.s - e.sExpect;
            e.logLike -= 0.5f * (ds * ds) * e.sExpectWeight;
            const float delta = e.logLike - e.logLikePrev;
            const float convTol = 1e-4f * hw_fabsf(e.logLike);
            bool conv = (delta >= 0.0f) && (delta < convTol);
            conv = conv && (e.logLike >= e.mlLogLike);
            // MaxLikelihoodDiagnostics::Converged (MaxLikelihoodDiagnostics.h:90-101)
            if (conv && !e.converged) e.iterCount = (int)it;
            if (conv) {
                e.mlLogLike = e.logLike;
                e.mlDelta = delta;
                e.captureIter = (int)it;
                for (int k = 0; k < 5; ++k) {
                    e.rhoEst[k] = e.rho[k];
                    e.muEst[k] = e.mu[k];
                    e.varEst[k] = e.var[k];
                }
            }
            e.converged = e.converged || conv;
            if (!(e.converged || e.status != HWPA_ZS_OK)) allDone = false;
        }
        // Lane-wide early exit (DmeEmHost.cpp:485)
        if (!d.iterToLimit && allDone) break;

        // M-step (DmeEmHost.cpp:494-564)
        for (int z = 0; z < LS; ++z) {
            EmZ& E = ez[z];
            const HistZ& H = hz[z];
            const float* tau = &tauBuf[(size_t)z * 5 * NE];
            const float* bp = &bpBuf[(size_t)z * NE];
            const float* binProb = &binProbBuf[(size_t)z * NB];
            float c[5];
            for (int m = 0; m < 5; ++m) c[m] = hw_maxf(ci[z][m], 0.1f);
            for (int m = 0; m < 5; ++m) E.rho[m] = c[m] / E.njSum;
            const float rs = Sum5Tree(E.rho);
            for (int m = 0; m < 5; ++m) E.rho[m] = E.rho[m] / rs;
            E.mu[0] = mom1[z][0] / c[0];
            float numer = 0.0f, denom = 0.0f;
            for (int a = 0; a < NA; ++a) {
                const int i = a + 1;
                numer += rpa[a] * varinv[z][i] * mom1[z][i];
                denom += (rpa[a] * rpa[a]) * varinv[z][i] * c[i];
            }
            numer += E.sExpect * E.sExpectWeight;
            denom += E.sExpectWeight;
            E.s = numer / denom;
            const float minPulseMean = hw_maxf(1.0f, d.k.signalScaler);
            const float sMin = hw_maxf(0.5f * hw_sqrtf(E.var[0]), minPulseMean) / rpaMin;
            if (E.s < sMin) {
                E.s = sMin;
                E.status |= HWPA_ZS_VLOW_SIGNAL;
            }
            for (int a = 0; a < NA; ++a) E.mu[a + 1] = E.s * rpa[a];
            float mom2 = 0.0f;
            for (int j = 0; j < NB; ++j) {
                const int k = j + 1;
                const float x0 = H.start[j] - E.mu[0];
                const float x1 = H.start[k] - E.mu[0];
                float tmp = bp[j] * x0 * x0 * tau[0 * NE + j];
                tmp += bp[k] * x1 * x1 * tau[0 * NE + k];
                tmp *= (float)H.cnt[j] / binProb[j];
                mom2 += tmp;
            }
            mom2 *= 0.5f * H.binSize;
            E.var[0] = mom2 / c[0] + E.varQuant;
            E.var[0] = hw_clampf(E.var[0], d.k.blVarMin, d.k.blVarMax);
            for (int a = 0; a < NA; ++a)
                E.var[a + 1] = ModelSignalCovar(d.k, d.cv[a] * d.cv[a], E.mu[a + 1], E.var[0]);
        }
    }
    return it;
}

// The learned initialiser's start for one lane: network outputs y [64][11] through
// the output map, the capped EM's limits and the acceptance threshold; `fallback`
// [64] receives 1 for a ZMW whose capped estimate was rejected (its model is then
// left as it was, for the plain EM).
struct LearnedStart {
    const float* y;
    const float* omap;
    unsigned iterLimit;
    float convTol, acceptConf, gateLogit;
    unsigned laneMinConf;
    uint8_t* fallback;
};

// DmeEmHost::EstimateLaneDetModel (DmeEmHost.cpp:216-615) for one lane.
template <class M>
void EstimateLane(const DmeCfg& d, const hwpa_lane_hist& lh, const hwpa_baseliner_stats& bs,
                  hwpa_lane_model& lm, const int32_t* modelFi, const int32_t* estFi, orc_dme_diag* dx,
                  const LearnedStart* ls = nullptr)
{
    static thread_local std::vector<EmZ> ez;
    ez.resize(LS);
    std::vector<HistZ> hz(LS);

    const float rpa[NA] = {d.relAmp[0], d.relAmp[1], d.relAmp[2], d.relAmp[3]};
    const float rpaSqSum = (rpa[0] * rpa[0] + rpa[2] * rpa[2]) + (rpa[1] * rpa[1] + rpa[3] * rpa[3]);
    const float rpaMin = std::min(std::min(rpa[0], rpa[1]), std::min(rpa[2], rpa[3]));

    const float evolve = EvolveFactor<M>(d, modelFi, estFi);

    for (int z = 0; z < LS; ++z) {
        EmZ& e = ez[z];
        HistZ& h = hz[z];
        MakeHistZ(lh, z, h);

        e.det = ModelFromHalf(lm, z);
        e.det.conf *= evolve;

        // PrelimEstimate (DmeEmHost.cpp:161-214)
        ModelZ w = e.det;
        {
            const float nBl = bs.baseline.m0[z];
            const float total = bs.full.basic.m0[z];
            const float blWeight = hw_maxf(nBl / total, 0.01f);
            const bool mask = nBl >= 3.0f;
            StatZ blsa{bs.baseline.offset[z], bs.baseline.m0[z], bs.baseline.m1[z], bs.baseline.m2[z]};
            float blVar = mask ? blsa.Variance() : w.bl.var;
            blVar = hw_clampf(blVar, d.k.blVarMin, d.k.blVarMax);
            const float blMean = mask ? blsa.Mean() : w.bl.mean;
            const float scale = hw_sqrtf(blVar / w.bl.var);
            for (int a = 0; a < NA; ++a) {
                const float mean = w.an[a].mean * scale;
                const float var = ModelSignalCovar(d.k, d.cv[a] * d.cv[a], mean, blVar);
                w.an[a].mean = mean;
                w.an[a].var = var;
                w.an[a].w = 0.25f * (1.0f - blWeight);
            }
            w.bl.w = blWeight;
            w.bl.mean = blMean;
            w.bl.var = blVar;
            float conf = 0.1f * hw_satlin(0.0f, 500.0f, nBl - 3.0f);
            conf *= hw_satlin(d.k.blVarMax, 0.5f * (d.k.blVarNominal + d.k.blVarMax), blVar);
            w.conf = conf;
        }
        UpdateConf<M>(d, e.det, w);
        e.work = e.det;

        e.numFrames = h.Total();
        e.scaleFactor = PrelimScaleFactor(d, e.work, h);
        // ScaleModelSnr (DmeEmHost.cpp:837-852)
        for (int a = 0; a < NA; ++a) {
            e.work.an[a].mean = e.scaleFactor * e.work.an[a].mean;
            e.work.an[a].var = ModelSignalCovar(d.k, d.cv[a] * d.cv[a], e.work.an[a].mean, e.work.bl.var);
        }
        const float rhoMin = 0.001f;
        e.rho[0] = hw_maxf(e.work.bl.w, rhoMin);
        e.mu[0] = e.work.bl.mean;
        e.var[0] = e.work.bl.var;
        for (int a = 0; a < NA; ++a) {
            e.rho[a + 1] = hw_maxf(e.work.an[a].w, rhoMin);
            e.mu[a + 1] = e.work.an[a].mean;
            e.var[a + 1] = e.work.an[a].var;
        }
        e.varQuant = (h.binSize * h.binSize) / 12.0f;
        for (int k = 0; k < 5; ++k) e.var[k] += e.varQuant;
        e.var[0] = hw_minf(e.var[0], d.k.blVarMax);
        const float rs = Sum5Tree(e.rho);
        for (int k = 0; k < 5; ++k) e.rho[k] = e.rho[k] / rs;
        for (int k = 0; k < 5; ++k) {
            e.rhoEst[k] = e.rho[k];
            e.muEst[k] = e.mu[k];
            e.varEst[k] = e.var[k];
        }
        e.s = ((e.mu[1] * rpa[0] + e.mu[2] * rpa[1]) + (e.mu[3] * rpa[2] + e.mu[4] * rpa[3])) / rpaSqSum;
        e.sExpect = e.s / e.scaleFactor;
        e.sExpectWeight = e.det.conf * d.pulseAmpReg;
        e.logLike = -FLT_MAX;
        e.logLikePrev = e.logLike;
        e.binProbMin = hw_maxf((float)e.numFrames * h.binSize, 1.0f) * FLT_MIN;
        e.status = (e.numFrames < 160) ? HWPA_ZS_INSUF_DATA : HWPA_ZS_OK;
        e.converged = false;
        e.iterCount = 0;
        e.captureIter = -1;
        e.mlLogLike = -INFINITY;
        e.mlDelta = -INFINITY;
        float njs = 0.0f;
        for (int b = 0; b < NB; ++b) njs += (float)h.cnt[b];
        e.njSum = njs;
        if (ls) {
            // The network's start (the device's EM_LEARNED start, hw_amd/csrc/dme.hip): output
            // j < 6 is off[j] + scale[j] y[j] in units of u = 2 binSize from the lower bound
            // (means) and log(var / u^2); outputs 6..10 are softmax mixture logits.
            const float* y = ls->y + (size_t)z * HWPA_MLP_OUT;
            const float u = 2.0f * h.binSize;
            float r[6];
            for (int j = 0; j < 6; ++j) r[j] = ls->omap[j] + ls->omap[6 + j] * y[j];
            float mx = y[6];
            for (int k = 1; k < 5; ++k) mx = hw_maxf(mx, y[6 + k]);
            float ex[5], se = 0.0f;
            for (int k = 0; k < 5; ++k) {
                ex[k] = hw_expf(y[6 + k] - mx);
                se += ex[k];
            }
            for (int k = 0; k < 5; ++k) e.rho[k] = hw_maxf(ex[k] / se, rhoMin);
            const float rs2 = Sum5Tree(e.rho);
            for (int k = 0; k < 5; ++k) e.rho[k] = e.rho[k] / rs2;
            e.mu[0] = lh.low_bound[z] + u * r[0];
            e.var[0] = hw_clampf((u * u) * hw_expf(r[1]), d.k.blVarMin, d.k.blVarMax);
            for (int a = 0; a < NA; ++a) {
                e.mu[a + 1] = lh.low_bound[z] + u * r[2 + a];
                e.var[a + 1] = ModelSignalCovar(d.k, d.cv[a] * d.cv[a], e.mu[a + 1], e.var[0]);
            }
            for (int k = 0; k < 5; ++k) {
                e.rhoEst[k] = e.rho[k];
                e.muEst[k] = e.mu[k];
                e.varEst[k] = e.var[k];
            }
            e.s = ((e.mu[1] * rpa[0] + e.mu[2] * rpa[1]) + (e.mu[3] * rpa[2] + e.mu[4] * rpa[3])) / rpaSqSum;
        }
    }

    // The learned start's confidence gate (the device's EM_LEARNED): a ZMW whose confidence
    // output is below the threshold takes no capped iteration of its own and is rejected; a
    // lane with fewer than laneMinConf ZMWs above it gates all of them; a lane with every ZMW
    // gated skips the capped EM.
    bool gated[LS] = {};
    unsigned iterLimit = d.iterLimit;
    if (ls) {
        unsigned nConf = 0;
        for (int z = 0; z < LS; ++z) {
            gated[z] = !(ls->y[(size_t)z * HWPA_MLP_OUT + 11] >= ls->gateLogit);
            nConf += gated[z] ? 0u : 1u;
        }
        if (nConf < ls->laneMinConf)
            for (int z = 0; z < LS; ++z) gated[z] = true;
        iterLimit = (nConf == 0 || nConf < ls->laneMinConf) ? 0u : ls->iterLimit;
    }
    const unsigned it = M::kReference ? EmIterationsRef<M>(d, ez, hz, rpa, rpaMin)
                                      : EmIterationsMirror(d, ez, hz, rpa, rpaMin, iterLimit, ls ? ls->convTol : 1e-4f,
                                                           ls ? gated : nullptr);
    for (int z = 0; z < LS; ++z) {
        EmZ& e = ez[z];
        if (!e.converged) e.status |= HWPA_ZS_NO_CONVERGE;
        e.work.bl.w = e.rhoEst[0];
        e.work.bl.mean = e.muEst[0];
        e.work.bl.var = e.varEst[0];
        for (int a = 0; a < NA; ++a) {
            e.work.an[a].w = e.rhoEst[a + 1];
            e.work.an[a].mean = e.muEst[a + 1];
            e.work.an[a].var = e.varEst[a + 1];
        }
        float cf[8];
        ComputeConfidence<M>(d, e.converged, e.det, e.work, cf);
        float conf = 1.0f;
        for (int k = 0; k < 8; ++k) conf *= cf[k];
        conf = hw_minf(hw_maxf(0.0f, conf), 1.0f);
        conf = (conf >= d.successThresh) ? conf : 0.0f;
        e.work.conf = conf;
        if (dx) {
            dx->converged[z] = e.converged;
            dx->iter_count[z] = e.iterCount;
            dx->capture_iter[z] = e.captureIter;
            dx->status[z] = e.status;
            dx->log_like[z] = e.mlLogLike;
            for (int k = 0; k < 8; ++k) dx->conf_factors[k][z] = cf[k];
            dx->scale_factor[z] = e.scaleFactor;
            dx->confidence[z] = conf;
            const ModeZ* modes[5] = {&e.work.bl, &e.work.an[0], &e.work.an[1], &e.work.an[2], &e.work.an[3]};
            for (int m = 0; m < 5; ++m) {
                dx->est_fp32[m][0][z] = modes[m]->w;
                dx->est_fp32[m][1][z] = modes[m]->mean;
                dx->est_fp32[m][2][z] = modes[m]->var;
            }
        }
        if (ls) {
            // Accept converged, status-OK estimates with enough confidence; the model of a
            // rejected ZMW stays as it was for the plain EM.
            const bool accept = !gated[z] && e.converged && e.status == HWPA_ZS_OK && conf >= ls->acceptConf;
            ls->fallback[z] = accept ? 0 : 1;
            if (!accept) continue;
        }
        UpdateConf<M>(d, e.det, e.work);
        if (dx) {
            const ModeZ* modes[5] = {&e.det.bl, &e.det.an[0], &e.det.an[1], &e.det.an[2], &e.det.an[3]};
            for (int m = 0; m < 5; ++m) {
                dx->out_fp32[m][0][z] = modes[m]->w;
                dx->out_fp32[m][1][z] = modes[m]->mean;
                dx->out_fp32[m][2][z] = modes[m]->var;
            }
        }
        ModelToHalf(e.det, lm, z);
    }
    if (dx) dx->iterations_run = (int)it;
}

}  // namespace

// ---------------------------------------------------------------------------
// C API
// ---------------------------------------------------------------------------

extern "C" {

int orc_threads(int n)
{
    const int prev = g_threads;
    if (n > 0) g_threads = n;
    return prev;
}

void orc_filter_params(int32_t type, hwpa_filter_params* out) { *out = FilterParams(type); }

void* orc_baseliner_create(const hwpa_config* cfg, int32_t lanes)
{
    auto* c = new BaselinerCtx;
    c->fp = FilterParams(cfg->filter_type);
    // LaneBaseliner ctor (HostMultiScaleBaseliner.h:58-66) and MultiScaleBaseliner::Configure
    // (MultiScaleBaseliner.cpp:53-75)
    c->cMeanBias = c->fp.mean_bias * std::exp2(cfg->mean_bias_adjust);
    c->cSigmaBias = c->fp.sigma_bias * std::exp2(cfg->sigma_bias_adjust);
    c->meanEmaAlpha = std::pow(0.5f, 1.0f / cfg->mean_ema_scale_strides);
    c->sigmaEmaAlpha = std::exp2(-1.0f / cfg->sigma_ema_scale_strides);
    c->jumpTol = (cfg->jump_suppression > 0.0f) ? 1.0f / cfg->jump_suppression : INFINITY;
    c->scaler = cfg->movie.photoelectron_sensitivity;
    c->pedestal = cfg->movie.pedestal;
    c->lanes = lanes;
    c->lb.resize(lanes);
    const auto& fp = c->fp;
    for (auto& L : c->lb) {
        // MultiStageFilter ctor (HostMultiScaleBaseliner.h:115-125)
        L.lower.resize(2 * fp.num_stages);
        L.upper.resize(2 * fp.num_stages);
        L.lower[0].Init(fp.widths[0], fp.strides[0], 1, false);
        L.lower[1].Init(fp.widths[0], 1, fp.strides[0], true);
        L.upper[0].Init(fp.widths[0], fp.strides[0], 1, true);
        L.upper[1].Init(fp.widths[0], 1, fp.strides[0], false);
        for (int i = 1; i < fp.num_stages; ++i) {
            for (auto* v : {&L.lower, &L.upper}) {
                (*v)[2 * i].Init(fp.widths[i], fp.strides[i], fp.strides[i - 1], false);
                (*v)[2 * i + 1].Init(fp.widths[i], 1, fp.strides[i - 1] * fp.strides[i], true);
            }
        }
        for (int z = 0; z < LS; ++z) {
            L.emaSum[z] = 0.0f;
            L.emaWeight[z] = 0.0f;
            L.sigmaEma[z] = 2.0f;
            L.latData[z] = L.latRaw[z] = 0;
            L.latL[z] = L.latH1[z] = L.latH2[z] = false;
        }
    }
    return c;
}

void orc_baseliner_destroy(void* h) { delete static_cast<BaselinerCtx*>(h); }

int32_t orc_baseliner_process(void* h, const int16_t* raw, int32_t frames, int16_t* out,
                              hwpa_baseliner_stats* stats)
{
    auto* c = static_cast<BaselinerCtx*>(h);
    if (frames <= 0 || frames % c->fp.aggregate_stride != 0) return HWPA_ERR_INVALID_ARG;
    ParLanes(c->lanes, [&](int l) {
        const size_t off = (size_t)l * frames * LS;
        EstimateBaselineLane(*c, l, raw + off, frames, out + off, stats + l);
    });
    return HWPA_OK;
}

struct AggCtx {
    int lanes;
    std::vector<LaneStats> st;
};

void* orc_agg_create(int32_t lanes)
{
    auto* a = new AggCtx;
    a->lanes = lanes;
    a->st.resize(lanes);
    return a;
}
void orc_agg_destroy(void* h) { delete static_cast<AggCtx*>(h); }

// BaselineStatsAggregatorHost::AddMetricsImpl (BaselineStatsAggregatorHost.cpp:40-51)
void orc_agg_add(void* h, const hwpa_baseliner_stats* stats)
{
    auto* a = static_cast<AggCtx*>(h);
    ParLanes(a->lanes, [&](int l) {
        LaneStats o;
        o.FromState(stats + l);
        a->st[l].Merge(o);
    });
}
void orc_agg_get(void* h, hwpa_baseliner_stats* out)
{
    auto* a = static_cast<AggCtx*>(h);
    for (int l = 0; l < a->lanes; ++l) a->st[l].ToState(out + l);
}
void orc_agg_reset(void* h)
{
    auto* a = static_cast<AggCtx*>(h);
    for (auto& s : a->st) s.Reset();
}

void* orc_hist_create(const hwpa_config* cfg, int32_t lanes)
{
    auto* c = new HistCtx;
    c->binSizeCoeff = cfg->bin_size_coeff;
    c->binSizeMin = cfg->bin_size_low_bound_coeff * std::max(1.0f, cfg->movie.photoelectron_sensitivity);
    c->fallbackSigma = cfg->fallback_baseline_sigma;
    c->minFrameCount = (float)cfg->baseline_stat_min_frame_count;
    c->lanes = lanes;
    c->h.resize(lanes);
    for (auto& H : c->h) {
        std::memset(H.prevFrame, 0, sizeof(H.prevFrame));
        std::memset(H.hist, 0, sizeof(H.hist));
        H.hi0 = 0;
        H.hi1 = -1;
        std::memset(H.count, 0, sizeof(H.count));
        for (int z = 0; z < LS; ++z) H.binSize[z] = 1.0f;
        std::memset(H.binStart, 0, sizeof(H.binStart));
        std::memset(H.lowOut, 0, sizeof(H.lowOut));
        std::memset(H.highOut, 0, sizeof(H.highOut));
    }
    return c;
}
void orc_hist_destroy(void* h) { delete static_cast<HistCtx*>(h); }

int32_t orc_hist_reset(void* h, const hwpa_baseliner_stats* stats)
{
    auto* c = static_cast<HistCtx*>(h);
    int err = 0;
    for (int l = 0; l < c->lanes; ++l) err |= HistResetLane(*c, l, stats[l]);
    return err ? HWPA_ERR_INVALID_ARG : HWPA_OK;
}

void orc_hist_add(void* h, const int16_t* traces, int32_t frames, const hwpa_lane_model* models)
{
    auto* c = static_cast<HistCtx*>(h);
    ParLanes(c->lanes, [&](int l) {
        HistAddLane(*c, l, traces + (size_t)l * frames * LS, frames, models[l]);
    });
}

// Loads bounds, counts and outliers of each lane (the LaneHistogram view of
// UHistogramSimd, dataTypes/UHistogramSimd.cpp:100-118); edge state is kept.
void orc_shortbin_sweep(const float* lo, const float* bs, int32_t n, uint64_t out[3])
{
    ShortbinSweep(lo, bs, n, out);
}

void orc_div_by_const_sweep(float d, const int32_t* exps, int32_t n, uint64_t out[2])
{
    const float rd = 1.0f / d;
    uint64_t checked = 0, wrong = 0;
    auto check = [&](float x) {
        const float q = x / d;
        const float c = hw_div_by_const(x, d, rd);
        const bool same = (q == c) || (hw_isnan(q) && hw_isnan(c));
        const bool sameSign = hw_isnan(q) || q == 0.0f || (hw_f2u(q) == hw_f2u(c));
        ++checked;
        wrong += (same && sameSign) ? 0u : 1u;
    };
    for (int32_t i = 0; i < n; ++i)
        for (uint32_t sign = 0; sign < 2; ++sign)
            for (uint32_t m = 0; m < (1u << 23); ++m)
                check(hw_u2f((sign << 31) | ((uint32_t)exps[i] << 23) | m));
    const uint32_t specials[] = {0x00000000u, 0x80000000u, 0x7f800000u, 0xff800000u, 0x7fc00000u, 0x7f7fffffu,
                                 0xff7fffffu};
    for (uint32_t u : specials) check(hw_u2f(u));
    out[0] = checked;
    out[1] = wrong;
}

void orc_roi_backward_words(const uint32_t* words, const uint32_t* carries, int32_t n, uint32_t* out_words,
                            uint32_t* out_carries)
{
    for (int32_t i = 0; i < n; ++i) {
        uint32_t c = carries[i];
        out_words[i] = hw_roi_backward_word(words[i], &c);
        out_carries[i] = c;
    }
}

void orc_hist_set(void* h, const hwpa_lane_hist* in)
{
    auto* c = static_cast<HistCtx*>(h);
    for (int l = 0; l < c->lanes; ++l) {
        LaneHistState& H = c->h[l];
        for (int z = 0; z < LS; ++z) {
            const float lo = in[l].low_bound[z], bs = in[l].bin_size[z];
            H.binSize[z] = bs;
            H.binStart[0][z] = lo;
            for (int i = 1; i <= NB; ++i) H.binStart[i][z] = bs * (float)i + lo;
            H.lowOut[z] = in[l].outlier_low[z];
            H.highOut[z] = in[l].outlier_high[z];
            for (int b = 0; b < NB; ++b) H.count[b][z] = in[l].bin_count[b][z];
            H.count[NB][z] = 0;
        }
    }
}

void orc_hist_get(void* h, hwpa_lane_hist* out)
{
    auto* c = static_cast<HistCtx*>(h);
    for (int l = 0; l < c->lanes; ++l) HistGetLane(*c, l, out[l]);
}

void orc_dme_init_models(const hwpa_config* cfg, int32_t lanes, const hwpa_baseliner_stats* stats,
                         hwpa_lane_model* out)
{
    const DmeCfg d = MakeDmeCfg(*cfg);
    for (int l = 0; l < lanes; ++l) InitLaneModel(d, stats[l], out[l]);
}

// Learned-start estimation as the device runs it (hwpa_dme_estimate_learned): the
// capped EM from the network outputs y [lanes * 64][11] with the output map omap
// ({offset[6], scale[6]}); then a lane whose 64 ZMWs were all rejected is re-estimated
// by the plain EM in place, and the rejected ZMWs of the other lanes are packed in lane
// order into compact lanes (the last one filled with copies of its first ZMW),
// re-estimated by the plain EM and written back.  Mirror arithmetic only (the path has no reference
// counterpart).  fallback [lanes * 64] (optional) receives the rejected ZMWs;
// returns their number.
int32_t orc_dme_estimate_learned(const hwpa_config* cfg, int32_t lanes, const hwpa_lane_hist* hist,
                                 const hwpa_baseliner_stats* stats, hwpa_lane_model* models, const int32_t model_fi[2],
                                 const int32_t est_fi[2], const float* y, const float* omap, uint8_t* fallback)
{
    const DmeCfg d = MakeDmeCfg(*cfg);
    std::vector<uint8_t> fb((size_t)lanes * LS, 0);
    ParLanes(lanes, [&](int l) {
        LearnedStart ls{y + (size_t)l * LS * HWPA_MLP_OUT, omap, cfg->learned_em_iteration_limit, cfg->learned_conv_tol,
                        cfg->learned_conf_threshold, hw_gate_logit(cfg->learned_gate_threshold),
                        cfg->learned_lane_min_confident, &fb[(size_t)l * LS]};
        EstimateLane<MirrorForm>(d, hist[l], stats[l], models[l], model_fi, est_fi, nullptr, &ls);
    });
    std::vector<uint32_t> map;
    std::vector<int> whole;
    for (int l = 0; l < lanes; ++l) {
        int n = 0;
        for (int z = 0; z < LS; ++z) n += fb[(size_t)l * LS + z] ? 1 : 0;
        if (n == LS) {
            whole.push_back(l);
            continue;
        }
        for (int z = 0; z < LS; ++z)
            if (fb[(size_t)l * LS + z]) map.push_back((uint32_t)(l * LS + z));
    }
    const uint32_t total = (uint32_t)map.size();
    const int32_t all = (int32_t)(total + (uint32_t)whole.size() * LS);
    if (fallback) std::memcpy(fallback, fb.data(), fb.size());
    ParLanes((int)whole.size(), [&](int i) {
        const int l = whole[i];
        EstimateLane<MirrorForm>(d, hist[l], stats[l], models[l], model_fi, est_fi, nullptr);
    });
    if (total == 0) return all;
    const int cl = (int)((total + LS - 1) / LS);
    std::vector<hwpa_lane_hist> ch(cl);
    std::vector<hwpa_baseliner_stats> cs(cl);
    std::vector<hwpa_lane_model> cm(cl);
    std::memset(ch.data(), 0, sizeof(hwpa_lane_hist) * cl);
    std::memset(cs.data(), 0, sizeof(hwpa_baseliner_stats) * cl);
    std::memset(cm.data(), 0, sizeof(hwpa_lane_model) * cl);
    auto copyModel = [](const hwpa_lane_model& m, int sz, hwpa_lane_model& dm, int dz) {
        for (int a = 0; a < NA; ++a) {
            dm.analogs[a].means[dz] = m.analogs[a].means[sz];
            dm.analogs[a].vars[dz] = m.analogs[a].vars[sz];
            dm.analogs[a].weights[dz] = m.analogs[a].weights[sz];
        }
        dm.baseline.means[dz] = m.baseline.means[sz];
        dm.baseline.vars[dz] = m.baseline.vars[sz];
        dm.baseline.weights[dz] = m.baseline.weights[sz];
        dm.confidence[dz] = m.confidence[sz];
    };
    auto copyZ = [&](const hwpa_lane_hist& h, const hwpa_baseliner_stats& st, const hwpa_lane_model& m, int sz,
                     hwpa_lane_hist& dh, hwpa_baseliner_stats& dst, hwpa_lane_model& dm, int dz) {
        dh.low_bound[dz] = h.low_bound[sz];
        dh.bin_size[dz] = h.bin_size[sz];
        dh.outlier_low[dz] = h.outlier_low[sz];
        dh.outlier_high[dz] = h.outlier_high[sz];
        for (int b = 0; b < NB; ++b) dh.bin_count[b][dz] = h.bin_count[b][sz];
        dst.baseline.offset[dz] = st.baseline.offset[sz];
        dst.baseline.m0[dz] = st.baseline.m0[sz];
        dst.baseline.m1[dz] = st.baseline.m1[sz];
        dst.baseline.m2[dz] = st.baseline.m2[sz];
        dst.full.basic.m0[dz] = st.full.basic.m0[sz];
        copyModel(m, sz, dm, dz);
    };
    for (uint32_t p = 0; p < total; ++p) {
        const uint32_t src = map[p];
        copyZ(hist[src / LS], stats[src / LS], models[src / LS], (int)(src % LS), ch[p / LS], cs[p / LS], cm[p / LS],
              (int)(p % LS));
    }
    for (uint32_t p = total; p < (uint32_t)cl * LS; ++p)
        copyZ(ch[cl - 1], cs[cl - 1], cm[cl - 1], 0, ch[cl - 1], cs[cl - 1], cm[cl - 1], (int)(p % LS));
    ParLanes(cl, [&](int l) { EstimateLane<MirrorForm>(d, ch[l], cs[l], cm[l], model_fi, est_fi, nullptr); });
    for (uint32_t p = 0; p < total; ++p) {
        const uint32_t dst = map[p];
        copyModel(cm[p / LS], (int)(p % LS), models[dst / LS], (int)(dst % LS));
    }
    return all;
}

void orc_dme_estimate(int32_t form, const hwpa_config* cfg, int32_t lanes, const hwpa_lane_hist* hist,
                      const hwpa_baseliner_stats* stats, hwpa_lane_model* models,
                      const int32_t model_fi[2], const int32_t est_fi[2], orc_dme_diag* diag)
{
    const DmeCfg d = MakeDmeCfg(*cfg);
    ParLanes(lanes, [&](int l) {
        orc_dme_diag* dx = diag ? diag + l : nullptr;
        if (form == ORC_FORM_MIRROR) EstimateLane<MirrorForm>(d, hist[l], stats[l], models[l], model_fi, est_fi, dx);
        else if (form == ORC_FORM_REFERENCE_LIBM)
            EstimateLane<ReferenceLibmForm>(d, hist[l], stats[l], models[l], model_fi, est_fi, dx);
        else EstimateLane<ReferenceForm>(d, hist[l], stats[l], models[l], model_fi, est_fi, dx);
    });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Frame labeler (traceAnalysis/FrameLabelerHost.cpp, SubframeScorer.*, RoiHost.h)
// ---------------------------------------------------------------------------

namespace {

// TransitionMatrix<float> ctor (traceAnalysis/SubframeScorer.cpp:35-223)
void BuildTransition(const hwpa_config& cfg, hwpa_transition& T)
{
    hwpa_analog analogs[NA];
    for (int i = 0; i < NA; ++i) analogs[i] = cfg.movie.analogs[i];
    float data[NS][NS];
    for (int i = 0; i < NS; ++i)
        for (int j = 0; j < NS; ++j) data[i][j] = 0.0f;
    for (int i = 0; i < NA; ++i) {
        const float delta = 0.01f;
        if (std::abs(analogs[i].ipd2_slow_step_ratio - 1.0f) < delta)
            analogs[i].pw2_slow_step_ratio = (analogs[i].pw2_slow_step_ratio < 1.0f) ? 1.0f - delta : 1.0f + delta;
        if (std::abs(analogs[i].ipd2_slow_step_ratio - 1.0f) < delta)
            analogs[i].ipd2_slow_step_ratio = (analogs[i].ipd2_slow_step_ratio < 1.0f) ? 1.0f - delta : 1.0f + delta;
    }
    float meanIpd = 0.0f;
    for (int i = 0; i < NA; ++i) meanIpd += analogs[i].inter_pulse_distance;
    meanIpd = meanIpd / NA;
    const double frameRate = cfg.movie.frame_rate;
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
    // Sparsity pattern (SubframeScorer.h:54-68): rows B and U: {B} + [TU..AD];
    // rows F and D: {F_i} + {U_i}.
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
    for (int i = 0; i < NS; ++i)
        for (int j = 0; j < NS; ++j) T.logp[i][j] = (data[i][j] > 0.0f) ? std::log(data[i][j]) : 0.0f;
}

struct Scorer {
    // BlockStateSubframeScorer<VF> per-ZMW terms (SubframeScorer.h:143-318)
    float bgMean, bInvVar, bgFixed, x0;
    float ffMean[NA], pInvVar[NA], ffFixed[NA], logPMean[NA], x1[NA];
    float dimMean, dimVar, dimNorm, dimScale;

    void Setup(const hwpa_lane_model& m, int z)
    {
        const float log2pi = 1.8378770664f;
        const float pi = 3.1415926536f;
        const float sqrtHalfPi = hw_sqrtf(0.5f * pi);
        const float bMean = hw_h2f(m.baseline.means[z]);
        const float bVar = hw_h2f(m.baseline.vars[z]);
        const float bSigma = hw_sqrtf(bVar);
        x0 = bMean + bSigma * sqrtHalfPi;
        bInvVar = 1.0f / bVar;
        for (int a = 0; a < NA; ++a) {
            const float aMean = hw_h2f(m.analogs[a].means[z]);
            logPMean[a] = hw_cephes_logf(aMean - bMean);
            const float aVar = hw_h2f(m.analogs[a].vars[z]);
            const float aSigma = hw_sqrtf(aVar);
            x1[a] = aMean - aSigma * sqrtHalfPi;
            pInvVar[a] = 1.0f / aVar;
        }
        {
            const float aMean = hw_h2f(m.analogs[3].means[z]);
            const float aVar = hw_h2f(m.analogs[3].vars[z]);
            float eMean = 0.5f * (bMean + aMean);
            float eVar = (1.0f / 6.0f) * (aMean - bMean);
            eVar += bVar + aVar;
            eVar *= 0.5f;
            dimMean = eMean;
            dimVar = eVar;
            dimNorm = -0.5f * (log2pi + hw_cephes_logf(eVar));
            dimScale = 0.5f / eVar;
        }
        const float normConst = log2pi / 2.0f;
        bgFixed = -0.5f * hw_cephes_logf(bVar) - normConst;
        bgMean = bMean;
        for (int a = 0; a < NA; ++a) {
            ffFixed[a] = -0.5f * hw_cephes_logf(hw_h2f(m.analogs[a].vars[z])) - normConst;
            ffMean[a] = hw_h2f(m.analogs[a].means[z]);
        }
    }

    float Score(float d, int state) const
    {
        if (state == 0) {
            const float y = d - bgMean;
            return -0.5f * y * bInvVar * y + bgFixed;
        }
        if (state <= 4) {
            const int i = state - 1;
            const float y = d - ffMean[i];
            return -0.5f * y * pInvVar[i] * y + ffFixed[i];
        }
        const int i = (state - 1) % 4;
        // SubframeScore
        const bool lowExt = d < x0;
        const float cap = lowExt ? x0 : x1[i];
        float fixedTerm = lowExt ? bInvVar : pInvVar[i];
        fixedTerm = -0.5f * fixedTerm;
        const bool ext = lowExt || (d > x1[i]);
        const float dc = d - cap;
        float s = (ext ? fixedTerm * (dc * dc) : 0.0f) - logPMean[i];
        if (i == NA - 1) {
            const float dm = d - dimMean;
            const float fb = dimNorm - (dm * dm) * dimScale;
            s = (x0 <= x1[i]) ? s : fb;
        }
        const float y = d - ffMean[i];
        const float fScore = -0.5f * y * pInvVar[i] * y + ffFixed[i];
        if (d > ffMean[i]) s = hw_minf(hw_minf(s, fScore - 1.0f), 0.0f);
        return s;
    }
};

// RoiFilterDefault::SmoothedVal, Roi::ForwardRecursion and Roi::BackwardRecursion
// (RoiHost.h:75-248, FrameLabelerHost.cpp:283-309) for one ZMW: seq(k), k in
// [0, nroi + 2), is the sigma-normalised trace and `before` the frame preceding it.
// Forward, frame k's smoothed value is the largest of x_k, (x_k-1 + x_k)/sqrt 2,
// (x_k-1 + x_k + x_k+1)/sqrt 3 and (x_k-1 + .. + x_k+2)/2; a frame is marked strong
// when that value reaches the upper threshold, or the lower one while the frame
// before was strong (hysteresis, the state carried in `boundary`), and weak when it
// reaches the lower threshold.  Backward, a frame is in the ROI when it is strong,
// or weak with the next frame in the ROI.  out[k * stride] receives 1 (ROI) or 0.
template <class Seq>
void RoiZ(float upper, float lower, float before, uint16_t boundary, Seq seq, int nroi, uint16_t* out,
          size_t stride)
{
    float prev[3] = {before, seq(0), seq(1)};
    uint16_t prevRoi = boundary;
    const float s2 = hw_sqrtf(2.0f), s3 = hw_sqrtf(3.0f);
    for (int k = 0; k < nroi; ++k) {
        const float nv = seq(k + 2);
        float mval = prev[1];
        mval = hw_maxf(mval, (prev[0] + prev[1]) / s2);
        mval = hw_maxf(mval, (prev[0] + prev[1] + prev[2]) / s3);
        mval = hw_maxf(mval, (prev[0] + prev[1] + prev[2] + nv) / 2.0f);
        const float thresh = (prevRoi & 1u) ? lower : upper;
        const uint16_t v = (uint16_t)((mval >= thresh ? 1u : 0u) | (mval >= lower ? 2u : 0u));
        out[(size_t)k * stride] = v;
        prevRoi = v;
        prev[0] = prev[1];
        prev[1] = prev[2];
        prev[2] = nv;
    }
    uint16_t pr = 0;
    for (int k = nroi - 1; k >= 0; --k) {
        const uint16_t r = out[(size_t)k * stride];
        const bool b = (r & 1u) || ((r & 2u) && pr != 0);
        pr = b ? 1 : 0;
        out[(size_t)k * stride] = b ? 1 : 0;
    }
}

struct LaneLabeler {
    hwpa_lane_model oldModel;
    int16_t labelBoundary[LS];
    uint16_t roiBoundary[LS];
    int16_t latentTrc[LS];
    int16_t prevLat[HWPA_LATENT_FRAMES][LS];
};

struct LabelerCtx {
    hwpa_transition T;
    float roiUpper, roiLower;
    int lanes;
    std::vector<LaneLabeler> L;
};

// LabelBlock (FrameLabelerHost.cpp:312-448) for one lane.
void LabelLane(const LabelerCtx& c, LaneLabeler& L, const int16_t* tr, int frames,
               const hwpa_lane_model& model, int16_t* labelsOut, int16_t* latOut, float* score)
{
    const int NL = HWPA_LATENT_FRAMES;
    const int nroi = frames + HWPA_VITERBI_LOOKBACK + HWPA_ROI_STITCH;
    const int nlab = frames + HWPA_VITERBI_LOOKBACK;
    std::vector<uint16_t> roi((size_t)nroi * LS);
    std::vector<uint8_t> tb((size_t)nlab * NS * LS);
    for (int i = 0; i < NL; ++i)
        for (int z = 0; z < LS; ++z) latOut[i * LS + z] = L.prevLat[i][z];

    for (int z = 0; z < LS; ++z) {
        // ComputeRoi (FrameLabelerHost.cpp:283-309, RoiHost.h:137-248)
        const float inv1 = 1.0f / hw_sqrtf(hw_h2f(L.oldModel.baseline.vars[z]));
        const float inv2 = 1.0f / hw_sqrtf(hw_h2f(model.baseline.vars[z]));
        auto seq = [&](int k) -> float {  // concatenated prevLat ++ traces, sigma-normalised
            return (k < NL) ? (float)L.prevLat[k][z] * inv1 : (float)tr[(size_t)(k - NL) * LS + z] * inv2;
        };
        RoiZ(c.roiUpper, c.roiLower, (float)L.latentTrc[z] * inv1, L.roiBoundary[z], seq, nroi, &roi[z], LS);

        // Viterbi recursion (FrameLabelerHost.cpp:213-264)
        float ll[NS];
        for (int s = 0; s < NS; ++s) ll[s] = (L.labelBoundary[z] == s) ? 0.0f : -INFINITY;
        Scorer sc;
        sc.Setup(L.oldModel, z);
        float anchorLL[NS];
        const int anchor = frames - NL;
        for (int f = 0; f < nlab; ++f) {
            if (f == NL) sc.Setup(model, z);
            const float d = (f < NL) ? (float)L.prevLat[f][z] : (float)tr[(size_t)(f - NL) * LS + z];
            const bool isRoi = roi[(size_t)f * LS + z] != 0;
            float acc[NS];
            for (int r = 0; r < NS; ++r) {
                const float sco = sc.Score(d, r);
                const int fc = c.T.first_col[r];
                float mx = sco + c.T.logp[r][fc] + ll[fc];
                int mi = fc;
                for (int p = c.T.seg_first[r]; p < c.T.seg_last[r]; ++p) {
                    const float v = sco + c.T.logp[r][p] + ll[p];
                    if (v >= mx) {
                        mx = v;
                        mi = p;
                    }
                }
                const float pen = (r == 0 || isRoi) ? 0.0f : INFINITY;
                acc[r] = mx - pen;
                tb[((size_t)f * NS + r) * LS + z] = (uint8_t)mi;
            }
            for (int r = 0; r < NS; ++r) ll[r] = acc[r];
            if (f == NL + anchor - 1) {
                for (int r = 0; r < NS; ++r) anchorLL[r] = ll[r];
            }
        }
        // Normalize (FrameLabelerHost.cpp:185-205)
        float prob[NS];
        float mx = ll[0];
        for (int i = 1; i < NS; ++i) mx = hw_maxf(ll[i], mx);
        float sum = 0.0f;
        for (int i = 0; i < NS; ++i) {
            prob[i] = hw_cephes_expf(ll[i] - mx);
            sum += prob[i];
        }
        const float inv = 1 / sum;
        for (int i = 0; i < NS; ++i) prob[i] *= inv;
        for (int i = nlab - 1; i > frames - 1; --i) {
            float np[NS];
            for (int s = 0; s < NS; ++s) np[s] = 0.0f;
            for (int from = 0; from < NS; ++from) {
                const int prevState = tb[((size_t)i * NS + from) * LS + z];
                for (int to = 0; to < NS; ++to) np[to] += (to == prevState) ? prob[from] : 0.0f;
            }
            for (int s = 0; s < NS; ++s) prob[s] = np[s];
        }
        float maxProb = prob[0];
        int anchorState = 0;
        for (int i = 1; i < NS; ++i) {
            const bool cond = maxProb > prob[i];
            maxProb = cond ? maxProb : prob[i];
            anchorState = cond ? anchorState : i;
        }
        int st = anchorState;
        for (int f = frames - 1; f >= 0; --f) {
            labelsOut[(size_t)f * LS + z] = (int16_t)st;
            st = tb[((size_t)f * NS + st) * LS + z];
        }
        score[z] = (frames >= NL) ? anchorLL[anchorState] : NAN;
        L.roiBoundary[z] = roi[(size_t)(frames - 1) * LS + z];
        L.labelBoundary[z] = (int16_t)anchorState;
        L.latentTrc[z] = tr[(size_t)(anchor - 1) * LS + z];
    }
    L.oldModel = model;
    for (int i = 0; i < NL; ++i)
        for (int z = 0; z < LS; ++z) L.prevLat[i][z] = tr[(size_t)(frames - NL + i) * LS + z];
}

// ---------------------------------------------------------------------------
// Pulse accumulator (traceAnalysis/HostPulseAccumulator.{h,cpp}, PulseAccumulator.cpp,
// dataTypes/Pulse.h)
// ---------------------------------------------------------------------------

struct Segment {
    uint32_t start[LS], end[LS];
    int16_t first[LS], last[LS], smax[LS];
    float total[LS], m2[LS];
    int16_t label[LS];
};

struct PulseCtx {
    float widthThresh, ampThresh;
    uint32_t maxCalls;
    uint8_t analogMap[NA];
    int lanes;
    std::vector<Segment> seg;
};

inline uint16_t ToFixed(float v) { return (uint16_t)(v / 0.1f + 0.5f); }

inline uint8_t NucleotideOf(const PulseCtx& c, int label)
{
    if (label > 2 * NA) label -= 2 * NA;
    if (label > NA) label -= NA;
    return (label >= 1 && label <= NA) ? c.analogMap[label - 1] : (uint8_t)4;
}

void PulseLane(PulseCtx& c, int l, const int16_t* labels, const int16_t* lat, const int16_t* cur, int frames,
               int32_t firstFrame, const hwpa_lane_model& model, hwpa_pulse* pulses, uint32_t* counts,
               uint32_t* ovf, hwpa_stat_acc& blst)
{
    Segment& S = c.seg[l];
    const int NL = HWPA_LATENT_FRAMES;
    const float maxSignal = 65535 * 0.1f;
    StatZ bls[LS];
    for (int z = 0; z < LS; ++z) {
        counts[z] = 0;
        ovf[z] = 0;
    }
    for (int z = 0; z < LS; ++z) {
        const float minMean = hw_h2f(model.analogs[NA - 1].means[z]);
        for (int rf = 0; rf < frames; ++rf) {
            const uint32_t absf = (uint32_t)(rf + firstFrame);
            const int16_t lab = labels[(size_t)rf * LS + z];
            const int16_t sig = (rf < NL) ? lat[(size_t)rf * LS + z] : cur[(size_t)(rf - NL) * LS + z];
            const bool isUp = (lab > NA) && (lab <= 2 * NA);
            const bool boundary = isUp || (lab == 0 && S.label[z] != 0);
            const bool isPulse = S.label[z] != 0;
            if (boundary && isPulse) {
                // LabelsSegment::ToPulse (HostPulseAccumulator.h:95-126)
                S.end[z] = absf;
                const int width = (int)(absf - S.start[z]);
                const float rawMean = (float)(S.total[z] + S.last[z] + S.first[z]) / width;
                const float rawMid = (float)S.total[z] / (width - 2);
                const bool keep = (width >= c.widthThresh) || (rawMean * width >= c.ampThresh * minMean);
                hwpa_pulse p{};
                p.start = S.start[z];
                p.width = (uint16_t)width;
                // std::min / std::max on floats (HostPulseAccumulator.h:95-126)
                auto smin = [](float a, float b) { return (b < a) ? b : a; };
                auto smax = [](float a, float b) { return (a < b) ? b : a; };
                p.mean_signal = ToFixed(smin(maxSignal, smax(0.0f, rawMean)));
                p.mid_signal = ToFixed(width < 3 ? 0.0f : smin(maxSignal, smax(0.0f, rawMid)));
                p.max_signal = ToFixed(smin(maxSignal, smax(0.0f, (float)S.smax[z])));
                p.signal_m2 = S.m2[z];
                p.label = NucleotideOf(c, S.label[z]);
                p.is_reject = !keep;
                // LaneVectorView::push_back (dataTypes/BatchVectors.h:105-118)
                if (counts[z] < c.maxCalls) pulses[(size_t)z * c.maxCalls + counts[z]++] = p;
                else ovf[z]++;
            }
            if (boundary) {
                // ResetSegment (HostPulseAccumulator.h:128-139)
                S.start[z] = absf;
                S.end[z] = 0;
                S.first[z] = sig;
                S.last[z] = 0;
                S.smax[z] = sig;
                S.total[z] = 0.0f;
                S.m2[z] = 0.0f;
                S.label[z] = lab;
            } else {
                // AddSignal (HostPulseAccumulator.h:141-148)
                S.total[z] = S.total[z] + S.last[z];
                S.m2[z] = S.m2[z] + (float)S.last[z] * S.last[z];
                S.last[z] = sig;
                S.smax[z] = std::max(S.smax[z], sig);
            }
            if (!isPulse && !boundary) bls[z].Add((float)sig);
        }
        blst.offset[z] = bls[z].off;
        blst.m0[z] = bls[z].m0;
        blst.m1[z] = bls[z].m1;
        blst.m2[z] = bls[z].m2;
    }
}

// ---------------------------------------------------------------------------
// Analyzer (analyzer/BatchAnalyzer.cpp:288-373, traceAnalysis/DetectionModelEstimator.cpp:93-160)
// ---------------------------------------------------------------------------

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

// ---------------------------------------------------------------------------
// HF metrics (traceAnalysis/HFMetricsFilterHost.cpp:37-116; the accumulator
// dataTypes/BasecallingMetricsAccumulator.cpp:62-504 and its trace metrics,
// dataTypes/TraceAnalysisMetrics.h:47-190)
// ---------------------------------------------------------------------------

constexpr uint8_t kLabelNone = 5;  // Pulse::NucleotideLabel::NONE (dataTypes/Pulse.h:52-57)

struct HfmTree {
    bool present = false;
    float maxHsw = 0, hsw[2] = {0, 0};
    std::vector<int16_t> left, right;
    std::vector<int8_t> feature, value;
    std::vector<float> threshold;
};

// StatAccumulator::Offset(value) (common/StatAccumulator.h:234-246), lane-wide early return.
static void StatSetOffset(StatZ* a, const float* value)
{
    bool same = true;
    for (int z = 0; z < LS; ++z) same = same && (value[z] == a[z].off);
    if (same) return;
    for (int z = 0; z < LS; ++z) {
        const float m1new = a[z].m1 + a[z].m0 * (a[z].off - value[z]);
        const float m2new = a[z].m2 + (m1new * m1new - a[z].m1 * a[z].m1) / a[z].m0;
        a[z].off = value[z];
        a[z].m1 = m1new;
        a[z].m2 = (a[z].m0 == 0.0f) ? 0.0f : m2new;
    }
}

// StatAccumulator::operator+= (common/StatAccumulator.h:145-170).
static void StatAddAssign(StatZ* a, StatZ* o)
{
    bool allZero = true, anyNan = false;
    for (int z = 0; z < LS; ++z) {
        allZero = allZero && (o[z].m0 == 0.0f);
        anyNan = anyNan || std::isnan(a[z].off);
    }
    if (allZero) return;
    if (anyNan) {
        for (int z = 0; z < LS; ++z)
            if (std::isnan(a[z].off)) a[z] = StatZ{0.0f, 0.0f, 0.0f, 0.0f};
    }
    float offNew[LS];
    for (int z = 0; z < LS; ++z) {
        const float w = o[z].m0 / (a[z].m0 + o[z].m0);
        const float on = (1.0f - w) * a[z].off + w * o[z].off;
        offNew[z] = std::isnan(on) ? a[z].off : on;
    }
    StatSetOffset(o, offNew);
    StatSetOffset(a, offNew);
    for (int z = 0; z < LS; ++z) {
        a[z].m0 += o[z].m0;
        a[z].m1 += o[z].m1;
        a[z].m2 += o[z].m2;
    }
}

// LaneArray sqrt (common/simd/m512f_SSE.h:406-415): 0 maps to 0.
static float LaneSqrt(float x) { return (x == 0.0f) ? 0.0f : hw_sqrtf(x); }

struct HfmLane {
    uint16_t numPulseFrames[LS], numBaseFrames[LS], numSandwiches[LS], numHalfSandwiches[LS], numStutters[LS];
    uint8_t activity[LS];
    float pkMidSignal[NA][LS], bpZvar[NA][LS], pkZvar[NA][LS], pkMax[NA][LS], modelVar[NA][LS], modelMean[NA][LS];
    uint16_t numPkMidFrames[NA][LS], numPkMidBases[NA][LS], numBasesBy[NA][LS], numPulsesBy[NA][LS];
    // TraceAnalysisMetrics
    uint32_t startFrame[LS], numFrames[LS];
    float pdScore[LS];
    StatZ blStat[LS];   // baselineStatAccum_: Reset keeps the offset
    AcLane ac;          // autocorrAccum_
    // basecall caches: [0] previous pulse, [1] the one before
    uint32_t cStart[2][LS];
    uint16_t cWidth[2][LS];
    uint8_t cLabel[2][LS];

    HfmLane()
    {
        for (int z = 0; z < LS; ++z) {
            startFrame[z] = numFrames[z] = 0;
            pdScore[z] = 0.0f;
            blStat[z] = StatZ{};
            for (int a = 0; a < NA; ++a) modelVar[a][z] = modelMean[a][z] = 0.0f;
        }
        Reset();
    }

    // BasecallingMetricsAccumulator::Reset (BasecallingMetricsAccumulator.cpp:254-277) and
    // TraceAnalysisMetrics::Reset (TraceAnalysisMetrics.h:71-84).
    void Reset()
    {
        for (int z = 0; z < LS; ++z) {
            numPulseFrames[z] = numBaseFrames[z] = numSandwiches[z] = numHalfSandwiches[z] = numStutters[z] = 0;
            activity[z] = 0;
            for (int c = 0; c < 2; ++c) {
                cStart[c][z] = 0;
                cWidth[c][z] = 0;
                cLabel[c][z] = kLabelNone;
            }
            for (int a = 0; a < NA; ++a) {
                pkMidSignal[a][z] = bpZvar[a][z] = pkZvar[a][z] = pkMax[a][z] = 0.0f;
                numPkMidFrames[a][z] = numPkMidBases[a][z] = numBasesBy[a][z] = numPulsesBy[a][z] = 0;
            }
            blStat[z].m0 = blStat[z].m1 = blStat[z].m2 = 0.0f;
            startFrame[z] = startFrame[z] + numFrames[z];
            numFrames[z] = 0;
            pdScore[z] = 0.0f;
        }
        ac.Reset();
    }

    // BasecallingMetricsAccumulator::Count (BasecallingMetricsAccumulator.cpp:333-410).  Labels
    // beyond T (never emitted by the pulse accumulator) are not counted per analog.
    void Count(const hwpa_pulse* pulses, const uint32_t* counts, uint32_t maxCalls, uint32_t frames)
    {
        for (int z = 0; z < LS; ++z) {
            numFrames[z] += frames;
            uint32_t pS = cStart[0][z], ppS = cStart[1][z];
            uint16_t pW = cWidth[0][z], ppW = cWidth[1][z];
            uint8_t pL = cLabel[0][z], ppL = cLabel[1][z];
            const uint32_t n = std::min(counts[z], maxCalls);
            for (uint32_t bi = 0; bi < n; ++bi) {
                const hwpa_pulse& p = pulses[(size_t)z * maxCalls + bi];
                const uint8_t label = p.label;
                numPulseFrames[z] = (uint16_t)(numPulseFrames[z] + p.width);
                if (label < NA) {
                    numPulsesBy[label][z] = (uint16_t)(numPulsesBy[label][z] + 1);
                    const float mx = (float)p.max_signal * 0.1f;
                    pkMax[label][z] = (pkMax[label][z] < mx) ? mx : pkMax[label][z];
                }
                if (pL != kLabelNone) {
                    if (pL == label) numStutters[z] = (uint16_t)(numStutters[z] + 1);
                    const bool abutted = p.start == pS + pW;
                    if (abutted && pL != label) numHalfSandwiches[z] = (uint16_t)(numHalfSandwiches[z] + 1);
                    if (ppL != kLabelNone) {
                        const bool prevAbutted = pS == ppS + ppW;
                        if (prevAbutted && abutted && ppL == label && ppL != pL)
                            numSandwiches[z] = (uint16_t)(numSandwiches[z] + 1);
                    }
                }
                if (!p.is_reject) {
                    numBaseFrames[z] = (uint16_t)(numBaseFrames[z] + p.width);
                    if (label < NA) {
                        numBasesBy[label][z] = (uint16_t)(numBasesBy[label][z] + 1);
                        if (p.width >= 3) {   // MidSignal() is NaN below 3 frames
                            numPkMidBases[label][z] = (uint16_t)(numPkMidBases[label][z] + 1);
                            const uint16_t midWidth = (uint16_t)(p.width - 2);
                            numPkMidFrames[label][z] = (uint16_t)(numPkMidFrames[label][z] + midWidth);
                            const float mid = (float)p.mid_signal * 0.1f;
                            pkMidSignal[label][z] += mid * (float)midWidth;
                            bpZvar[label][z] += mid * mid * (float)midWidth;
                            pkZvar[label][z] += p.signal_m2;
                        }
                    }
                }
                ppS = pS;
                ppW = pW;
                ppL = pL;
                pS = p.start;
                pW = p.width;
                pL = label;
            }
            cStart[0][z] = pS;
            cWidth[0][z] = pW;
            cLabel[0][z] = pL;
            cStart[1][z] = ppS;
            cWidth[1][z] = ppW;
            cLabel[1][z] = ppL;
        }
    }

    // BasecallingMetricsAccumulator::AddModels (BasecallingMetricsAccumulator.cpp:302-313)
    void AddModels(const hwpa_lane_model& m)
    {
        for (int a = 0; a < NA; ++a)
            for (int z = 0; z < LS; ++z) {
                modelVar[a][z] = hw_h2f(m.analogs[a].vars[z]);
                modelMean[a][z] = hw_h2f(m.analogs[a].means[z]);
            }
    }

    // BasecallingMetricsAccumulator::AddBatchMetrics (BasecallingMetricsAccumulator.cpp:315-331)
    void AddBatchMetrics(const hwpa_baseliner_stats& bs, const float* score, const hwpa_stat_acc& pd)
    {
        AcLane o;
        o.FromState(bs.full);
        ac.Merge(o);
        StatZ p[LS];
        for (int z = 0; z < LS; ++z) {
            pdScore[z] += score[z];
            const float bgMean = bs.background.m1[z] / bs.background.m0[z] + bs.background.offset[z];
            p[z] = StatZ{pd.offset[z] + bgMean, pd.m0[z], pd.m1[z], pd.m2[z]};
        }
        StatAddAssign(blStat, p);
    }

    uint16_t NumPulses(int z) const
    {
        uint16_t r = 0;
        for (int a = 0; a < NA; ++a) r = (uint16_t)(r + numPulsesBy[a][z]);
        return r;
    }
    uint16_t NumBases(int z) const
    {
        uint16_t r = 0;
        for (int a = 0; a < NA; ++a) r = (uint16_t)(r + numBasesBy[a][z]);
        return r;
    }

    // BasecallingMetricsAccumulator::FinalizeMetrics (BasecallingMetricsAccumulator.cpp:412-500)
    void Finalize(bool realtime, float frameRate, const HfmTree& tree)
    {
        for (int a = 0; a < NA; ++a) {
            bool allZero = true, allFewBases = true, allFewFrames = true;
            for (int z = 0; z < LS; ++z) {
                allZero = allZero && ((int16_t)numPkMidBases[a][z] == 0);
                allFewBases = allFewBases && ((float)numPkMidBases[a][z] < 2.0f);
                allFewFrames = allFewFrames && ((float)numPkMidFrames[a][z] < 2.0f);
            }
            if (allZero) {
                for (int z = 0; z < LS; ++z) pkMidSignal[a][z] = bpZvar[a][z] = pkZvar[a][z] = NAN;
            } else if (allFewBases || allFewFrames) {
                for (int z = 0; z < LS; ++z) bpZvar[a][z] = pkZvar[a][z] = NAN;
            } else {
                for (int z = 0; z < LS; ++z) {
                    float bpZ = bpZvar[a][z], pkZ = pkZvar[a][z], pkM = pkMidSignal[a][z];
                    const float nf = (float)numPkMidFrames[a][z];
                    const float nb = (float)numPkMidBases[a][z];
                    bpZ = (bpZ - (pkM * pkM / nf)) / nf;
                    bpZ = bpZ * nb / (nb - 1.0f);
                    const float blVar = blStat[z].Variance();
                    bpZ -= blVar / (nf / nb);
                    pkZ = (pkZ - (pkM * pkM / nf)) / (nf - 1.0f);
                    pkZ -= bpZ + blVar;
                    const float miv = modelVar[a][z];
                    bool mask = miv < blVar;
                    pkZ /= miv - blVar;
                    mask = mask || (pkZ < 0.0f);
                    pkZ = mask ? 0.0f : pkZ;
                    const float pkMid = pkM / nf;
                    mask = pkMid < 0.0f;
                    bpZ /= (pkMid * pkMid);
                    mask = mask || (bpZ < 0.0f);
                    bpZ = mask ? NAN : bpZ;
                    const int16_t nbi = (int16_t)numPkMidBases[a][z];
                    mask = nbi == 0;
                    pkM = mask ? NAN : pkM;
                    mask = mask || (nbi < 2);
                    mask = mask || ((int16_t)numPkMidFrames[a][z] < 2);
                    bpZ = mask ? NAN : bpZ;
                    pkZ = mask ? NAN : pkZ;
                    pkMidSignal[a][z] = pkM;
                    bpZvar[a][z] = bpZ;
                    pkZvar[a][z] = pkZ;
                }
            }
        }
        for (int z = 0; z < LS; ++z) {
            const float pd = pdScore[z] / (float)numFrames[z];
            pdScore[z] = std::isnan(pd) ? 0.0f : pd;
        }
        if (realtime && tree.present) LabelBlock(frameRate, tree);
    }

    // BasecallingMetricsAccumulator::LabelBlock (BasecallingMetricsAccumulator.cpp:62-200)
    void LabelBlock(float frameRate, const HfmTree& t)
    {
        auto z0 = [](float v) { return std::isnan(v) ? 0.0f : v; };
        for (int z = 0; z < LS; ++z) {
            float feat[11] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
            const float stdDev = LaneSqrt(blStat[z].Variance());
            const float nBases = (float)NumBases(z);
            const float nPulses = (float)NumPulses(z);
            const float pulseWidth = z0((float)numPulseFrames[z] / nPulses);
            float pkmid[NA];
            for (int a = 0; a < NA; ++a) pkmid[a] = pkMidSignal[a][z] / (float)numPkMidFrames[a][z];
            const float seconds = (float)numFrames[z] / frameRate;
            feat[0] = nPulses / seconds;
            feat[1] = z0((float)numSandwiches[z] / nPulses);
            const float hswr = z0((float)numHalfSandwiches[z] / nPulses);
            float hswrExp = t.hsw[0] * feat[0] + t.hsw[1];
            hswrExp = (hswrExp > t.maxHsw) ? t.maxHsw : hswrExp;
            feat[2] = hswr - hswrExp;
            feat[3] = pdScore[z];
            feat[4] = pulseWidth;
            feat[5] = z0((float)numStutters[z] / nPulses);
            float rel[NA];
            for (int a = 0; a < NA; ++a) rel[a] = modelMean[a][z];
            float maxamp = rel[0];
            for (int a = 1; a < NA; ++a) maxamp = hw_maxf(rel[a], maxamp);
            for (int a = 0; a < NA; ++a) rel[a] /= maxamp;
            int lowIdx = 0;
            float minamp = rel[0];
            for (int a = 1; a < NA; ++a) {
                lowIdx = (rel[a] < minamp) ? a : lowIdx;
                minamp = hw_minf(rel[a], minamp);
            }
            for (int a = 0; a < NA; ++a)
                feat[6] += pkmid[a] / stdDev * (float)numBasesBy[a][z] / nBases * minamp / rel[a];
            feat[6] = z0(feat[6]);
            for (int a = 0; a < NA; ++a) feat[7] = hw_maxf(feat[7], (pkMax[a][z] - pkmid[a]) / stdDev);
            feat[7] = z0(feat[7]);
            feat[8] = ac.Autocorrelation(z);
            for (int a = 0; a < NA; ++a) {
                const float b = bpZvar[a][z], k = pkZvar[a][z];
                feat[9] += (std::isnan(b) || lowIdx == a) ? 0.0f : b;
                feat[10] += (std::isnan(k) || lowIdx == a) ? 0.0f : k;
            }
            feat[9] /= 3.0f;
            feat[10] /= 3.0f;
            int cur = 0;
            for (size_t guard = 0; guard < t.feature.size() && t.feature[cur] >= 0; ++guard)
                cur = (feat[t.feature[cur]] <= t.threshold[cur]) ? t.left[cur] : t.right[cur];
            activity[z] = (uint8_t)t.value[cur];
        }
    }

    // BasecallingMetricsAccumulator::PopulateBasecallingMetrics (BasecallingMetricsAccumulator.cpp:202-252)
    void Populate(hwpa_basecalling_metrics& m) const
    {
        for (int z = 0; z < LS; ++z) {
            const float acr = ac.Autocorrelation(z);
            m.activity_label[z] = activity[z];
            m.num_pulse_frames[z] = numPulseFrames[z];
            m.num_base_frames[z] = numBaseFrames[z];
            m.num_sandwiches[z] = numSandwiches[z];
            m.num_half_sandwiches[z] = numHalfSandwiches[z];
            m.num_pulse_label_stutters[z] = numStutters[z];
            m.num_bases[z] = NumBases(z);
            m.num_pulses[z] = NumPulses(z);
            for (int a = 0; a < NA; ++a) {
                m.pk_mid_signal[a][z] = pkMidSignal[a][z];
                m.bp_zvar[a][z] = bpZvar[a][z];
                m.pk_zvar[a][z] = pkZvar[a][z];
                m.pk_max[a][z] = pkMax[a][z];
                m.num_pk_mid_frames[a][z] = numPkMidFrames[a][z];
                m.num_pk_mid_bases_by_analog[a][z] = numPkMidBases[a][z];
                m.num_bases_by_analog[a][z] = numBasesBy[a][z];
                m.num_pulses_by_analog[a][z] = numPulsesBy[a][z];
            }
            m.frame_baseline_dws[z] = blStat[z].Mean();
            m.frame_baseline_variance_dws[z] = blStat[z].Variance();
            m.num_frames_baseline[z] = (uint16_t)hw_round_i32(blStat[z].m0);
            m.start_frame[z] = startFrame[z];
            m.num_frames[z] = (uint16_t)numFrames[z];
            m.autocorrelation[z] = std::isnan(acr) ? 0.0f : acr;
            m.pulse_detection_score[z] = pdScore[z];
            m.pixel_checksum[z] = 0;
        }
    }
};

struct OrcMetrics {
    hwpa_config cfg;
    int lanes = 0;
    uint32_t framesSeen = 0;
    std::vector<HfmLane> lane;
    HfmTree tree;
};

enum { ST_HIST_INIT = 0, ST_DME_INIT = 1, ST_SEQUENCING = 2 };

}  // namespace

struct OrcAnalyzer {
    hwpa_config cfg;
    uint32_t poolId;
    int lanes, frames;
    int32_t nextFrame = 0;
    uint32_t baselinerLatency, dmeDelay;
    int32_t emForm = ORC_FORM_REFERENCE;
    bool fullEstimation = false;
    bool modelInitialized = false;
    int status = ST_HIST_INIT;
    int32_t framesRemaining;
    void* bl;
    void* agg;
    void* hist;
    void* lab;
    void* pulse;
    void* metrics;
    Interval aggFi, histFi, modelFi;
    std::vector<hwpa_lane_model> models;
    std::vector<int16_t> baselined, labels, latent;
    std::vector<hwpa_baseliner_stats> stats, aggStats;
    std::vector<hwpa_lane_hist> hists;
    std::vector<float> score;
};

extern "C" {

void orc_transition_matrix(const hwpa_config* cfg, hwpa_transition* out) { BuildTransition(*cfg, *out); }

void* orc_labeler_create(const hwpa_config* cfg, int32_t lanes)
{
    auto* c = new LabelerCtx;
    BuildTransition(*cfg, c->T);
    c->roiUpper = cfg->roi_upper_threshold;
    c->roiLower = cfg->roi_lower_threshold;
    c->lanes = lanes;
    c->L.resize(lanes);
    for (auto& L : c->L) {
        // LatentViterbi ctor (FrameLabelerHost.cpp:58-81); weights/confidence unused
        std::memset(&L.oldModel, 0, sizeof(L.oldModel));
        for (int z = 0; z < LS; ++z) {
            L.oldModel.baseline.means[z] = hw_f2h(0.0f);
            L.oldModel.baseline.vars[z] = hw_f2h(50.0f);
            for (int a = 0; a < NA; ++a) {
                L.oldModel.analogs[a].means[z] = hw_f2h(1000.0f);
                L.oldModel.analogs[a].vars[z] = hw_f2h(10.0f);
            }
            L.labelBoundary[z] = 0;
            L.roiBoundary[z] = 0;
            L.latentTrc[z] = 0;
        }
        std::memset(L.prevLat, 0, sizeof(L.prevLat));
    }
    return c;
}
void orc_labeler_destroy(void* h) { delete static_cast<LabelerCtx*>(h); }

void orc_labeler_process(void* h, const int16_t* traces, int32_t frames, const hwpa_lane_model* models,
                         int16_t* labels, int16_t* latent_trace, float* scores)
{
    auto* c = static_cast<LabelerCtx*>(h);
    ParLanes(c->lanes, [&](int l) {
        const size_t off = (size_t)l * frames * LS;
        LabelLane(*c, c->L[l], traces + off, frames, models[l], labels + off,
                  latent_trace + (size_t)l * HWPA_LATENT_FRAMES * LS, scores + (size_t)l * LS);
    });
}

void* orc_pulse_create(const hwpa_config* cfg, int32_t lanes)
{
    auto* c = new PulseCtx;
    c->widthThresh = cfg->xsp_width_thresh;
    const double x = cfg->xsp_amp_thresh;
    c->ampThresh = (float)(x / (1.0 - x));
    c->maxCalls = cfg->max_calls_per_zmw;
    for (int a = 0; a < NA; ++a) {
        const uint8_t b = cfg->movie.analogs[a].base_label;
        c->analogMap[a] = (b == 'A') ? 0 : (b == 'C') ? 1 : (b == 'G') ? 2 : (b == 'T') ? 3 : 4;
    }
    c->lanes = lanes;
    c->seg.resize(lanes);
    for (auto& S : c->seg) {
        std::memset(&S, 0, sizeof(S));  // default LabelsSegment: baseline label 0
    }
    return c;
}
void orc_pulse_destroy(void* h) { delete static_cast<PulseCtx*>(h); }

void orc_pulse_process(void* h, const int16_t* labels, const int16_t* latent_trace, const int16_t* cur_trace,
                       int32_t frames, int32_t first_frame, const hwpa_lane_model* models, hwpa_pulse* pulses,
                       uint32_t* counts, uint32_t* overflows, hwpa_stat_acc* bl_stats)
{
    auto* c = static_cast<PulseCtx*>(h);
    ParLanes(c->lanes, [&](int l) {
        const size_t off = (size_t)l * frames * LS;
        PulseLane(*c, l, labels + off, latent_trace + (size_t)l * HWPA_LATENT_FRAMES * LS, cur_trace + off,
                  frames, first_frame, models[l], pulses + (size_t)l * LS * c->maxCalls, counts + (size_t)l * LS,
                  overflows + (size_t)l * LS, bl_stats[l]);
    });
}

void orc_tracegen(const hwpa_tracegen_params* p, int32_t lanes, int32_t lane_offset, int32_t frames,
                  uint32_t first_frame, hwpa_tracegen_state* state, int32_t init, int16_t* out, int16_t* truth)
{
    ParLanes(lanes, [&](int l) {
        hwpa_tracegen_state& st = state[l];
        for (int i = 0; i < LS; ++i) {
            const uint64_t z = (uint64_t)(l + lane_offset) * LS + (uint64_t)i;
            if (init) hw_tracegen_init_zmw(p, z, &st, i);
            for (int t = 0; t < frames; ++t) {
                const size_t o = ((size_t)l * frames + t) * LS + i;
                if (truth) truth[o] = st.in_pulse[i] ? (int16_t)(st.analog[i] + 1) : (int16_t)0;
                out[o] = hw_tracegen_frame(p, z, first_frame + (uint32_t)t, &st, i);
            }
        }
    });
}

void* orc_metrics_create(const hwpa_config* cfg, int32_t lanes)
{
    auto* m = new OrcMetrics;
    m->cfg = *cfg;
    m->lanes = lanes;
    m->lane.resize(lanes);
    return m;
}

void orc_metrics_destroy(void* h) { delete static_cast<OrcMetrics*>(h); }

void orc_metrics_set_tree(void* h, const hwpa_activity_tree* t)
{
    auto* m = static_cast<OrcMetrics*>(h);
    HfmTree& r = m->tree;
    r = HfmTree{};
    if (!t || t->num_nodes <= 0) return;
    const int n = t->num_nodes;
    r.present = true;
    r.maxHsw = t->max_acceptable_hsw_rate;
    r.hsw[0] = t->hsw_curve[0];
    r.hsw[1] = t->hsw_curve[1];
    r.left.assign(t->children_left, t->children_left + n);
    r.right.assign(t->children_right, t->children_right + n);
    r.feature.assign(t->feature, t->feature + n);
    r.threshold.assign(t->threshold, t->threshold + n);
    r.value.assign(t->value, t->value + n);
}

// HFMetricsFilterHost::Process (HFMetricsFilterHost.cpp:86-116) for one chunk of pulses whose
// batch ends at last_frame; returns 1 and fills out[lanes] when a metrics block ends.
int32_t orc_metrics_process(void* h, const hwpa_pulse* pulses, const uint32_t* counts, uint32_t max_calls,
                            int32_t frames, int32_t last_frame, const hwpa_baseliner_stats* bl_stats,
                            const float* viterbi_score, const hwpa_stat_acc* pd_stats,
                            const hwpa_lane_model* models, hwpa_basecalling_metrics* out)
{
    auto* m = static_cast<OrcMetrics*>(h);
    if (m->framesSeen == 0)
        for (auto& l : m->lane) l.Reset();
    for (int l = 0; l < m->lanes; ++l) {
        HfmLane& a = m->lane[l];
        a.Count(pulses + (size_t)l * LS * max_calls, counts + (size_t)l * LS, max_calls, (uint32_t)frames);
        a.AddModels(models[l]);
        a.AddBatchMetrics(bl_stats[l], viterbi_score + (size_t)l * LS, pd_stats[l]);
    }
    m->framesSeen += (uint32_t)frames;
    if ((uint32_t)last_frame % m->cfg.frames_per_hf_metric_block >= (uint32_t)frames) return 0;
    for (int l = 0; l < m->lanes; ++l) {
        m->lane[l].Finalize(m->cfg.realtime_activity_labels != 0, m->cfg.movie.frame_rate, m->tree);
        if (out) m->lane[l].Populate(out[l]);
    }
    m->framesSeen = 0;
    return 1;
}

// Autocorrelation of the series x[0..n) added one value at a time (every ZMW the same) in
// `pieces` consecutive accumulators of sizes[p] values (equal sizes when sizes is null)
// merged left to right.
float orc_autocorr(const float* x, int32_t n, const int32_t* sizes, int32_t pieces)
{
    AcLane total;
    const int per = (n + pieces - 1) / pieces;
    int begin = 0;
    for (int p = 0; p < pieces; ++p) {
        AcLane part;
        const int end = sizes ? std::min(n, begin + sizes[p]) : std::min(n, (p + 1) * per);
        for (int i = begin; i < end; ++i) {
            float v[LS];
            for (int z = 0; z < LS; ++z) v[z] = x[i];
            part.AddSample(v);
        }
        begin = end;
        if (p == 0) total = part;
        else total.Merge(part);
    }
    return total.Autocorrelation(0);
}

void* orc_analyzer_create(const hwpa_config* cfg, uint32_t pool_id, int32_t lanes, int32_t frames)
{
    auto* a = new OrcAnalyzer;
    a->cfg = *cfg;
    a->poolId = pool_id;
    a->lanes = lanes;
    a->frames = frames;
    const hwpa_filter_params fp = FilterParams(cfg->filter_type);
    // MultiScaleBaseliner::StartupLatency (MultiScaleBaseliner.h:56-99)
    const float meanAlpha = std::pow(0.5f, 1.0f / cfg->mean_ema_scale_strides);
    const float sigmaAlpha = std::exp2(-1.0f / cfg->sigma_ema_scale_strides);
    const float mess = -1.0f / std::log2(meanAlpha);
    const float sess = -1.0f / std::log2(sigmaAlpha);
    a->baselinerLatency = (uint32_t)((float)fp.latent_size + std::max(mess, sess) * (float)fp.aggregate_stride);
    // DynamicEstimateBatchAnalyzer ctor stagger (BatchAnalyzer.cpp:100-118)
    const uint32_t cpe = (std::max(cfg->min_frames_for_estimate, 1u) + frames - 1) / frames;
    a->dmeDelay = (pool_id % cpe) * (uint32_t)frames;
    a->framesRemaining = (int32_t)cfg->num_frames_pre_accum_stats;
    a->bl = orc_baseliner_create(cfg, lanes);
    a->agg = orc_agg_create(lanes);
    a->hist = orc_hist_create(cfg, lanes);
    a->lab = orc_labeler_create(cfg, lanes);
    a->pulse = orc_pulse_create(cfg, lanes);
    a->metrics = orc_metrics_create(cfg, lanes);
    a->models.resize(lanes);
    std::memset(a->models.data(), 0, sizeof(hwpa_lane_model) * lanes);
    if (cfg->model_estimation_mode == HWPA_MODEL_FIXED) {
        // FixedModelBatchAnalyzer ctor (analyzer/BatchAnalyzer.cpp:120-145) with
        // StaticDetModelConfig::SetupAnalogs (dataTypes/configs/StaticDetModelConfig.h:53-68).
        hwpa_lane_model& m = a->models[0];
        const float blMean = cfg->static_baseline_mean, blVar = cfg->static_baseline_variance;
        const float refSignal = cfg->movie.ref_snr * std::sqrt(blVar);
        for (int z = 0; z < LS; ++z) {
            m.baseline.means[z] = hw_f2h(blMean);
            m.baseline.vars[z] = hw_f2h(blVar);
        }
        for (int an = 0; an < NA; ++an) {
            const float mean = blMean + cfg->movie.analogs[an].rel_amplitude * refSignal;
            const float cvm = cfg->movie.analogs[an].excess_noise_cv * mean;
            const float var = blVar + mean + cvm * cvm;   // std::pow(cv * mean, 2.f)
            for (int z = 0; z < LS; ++z) {
                m.analogs[an].means[z] = hw_f2h(mean);
                m.analogs[an].vars[z] = hw_f2h(var);
            }
        }
        for (int l = 1; l < lanes; ++l) a->models[l] = m;
    }
    a->baselined.resize((size_t)lanes * frames * LS);
    a->labels.resize((size_t)lanes * frames * LS);
    a->latent.resize((size_t)lanes * HWPA_LATENT_FRAMES * LS);
    a->stats.resize(lanes);
    a->aggStats.resize(lanes);
    a->hists.resize(lanes);
    a->score.resize((size_t)lanes * LS);
    return a;
}

void orc_analyzer_destroy(void* h)
{
    auto* a = static_cast<OrcAnalyzer*>(h);
    orc_baseliner_destroy(a->bl);
    orc_agg_destroy(a->agg);
    orc_hist_destroy(a->hist);
    orc_labeler_destroy(a->lab);
    orc_pulse_destroy(a->pulse);
    orc_metrics_destroy(a->metrics);
    delete a;
}

// DetectionModelEstimator::AddBatch (traceAnalysis/DetectionModelEstimator.cpp:93-160);
// *ran tells whether an estimate ran.
static int32_t OrcDmeAddBatch(OrcAnalyzer* a, int32_t first_frame, bool* ran)
{
    const int L = a->lanes, F = a->frames;
    *ran = false;
    orc_agg_add(a->agg, a->stats.data());
    a->aggFi.Hull(first_frame, first_frame + F);
    if (a->status != ST_HIST_INIT) {
        orc_hist_add(a->hist, a->baselined.data(), F, a->models.data());
        a->histFi.Hull(first_frame, first_frame + F);
    }
    if (a->status != ST_SEQUENCING) {
        orc_agg_get(a->agg, a->aggStats.data());
        orc_dme_init_models(&a->cfg, L, a->aggStats.data(), a->models.data());
        a->modelFi = a->aggFi;
    }
    a->framesRemaining -= F;
    if (a->framesRemaining <= 0) {
        if (a->status == ST_DME_INIT) a->status = ST_SEQUENCING;
        if (a->status == ST_SEQUENCING) {
            *ran = true;
            orc_hist_get(a->hist, a->hists.data());
            orc_agg_get(a->agg, a->aggStats.data());
            const int32_t mfi[2] = {a->modelFi.lo, a->modelFi.hi};
            const int32_t efi[2] = {a->histFi.lo, a->histFi.hi};
            orc_dme_estimate(a->emForm, &a->cfg, L, a->hists.data(), a->aggStats.data(), a->models.data(), mfi, efi,
                             nullptr);
            a->modelFi = a->histFi;
        }
        orc_agg_get(a->agg, a->aggStats.data());
        const int32_t rc = orc_hist_reset(a->hist, a->aggStats.data());
        if (rc != HWPA_OK) return rc;
        a->histFi = Interval{};
        orc_agg_reset(a->agg);
        a->aggFi = Interval{};
        if (a->status == ST_HIST_INIT) a->status = ST_DME_INIT;
        a->framesRemaining = (int32_t)a->cfg.min_frames_for_estimate;
    }
    return HWPA_OK;
}

// BatchAnalyzer::operator() (analyzer/BatchAnalyzer.cpp:149-170) and the AnalyzeImpl of the
// configured mode: FixedModel (:172-205), SingleEstimate (:207-283), DynamicEstimate (:288-373).
// uint8 samples are zero-extended to int16 first (dataTypes/BatchData.h:160-182).
int32_t orc_analyzer_process_encoded(void* h, const void* traces, int32_t encoding, uint32_t pool_id,
                                     int32_t first_frame, orc_result* res)
{
    auto* a = static_cast<OrcAnalyzer*>(h);
    if (encoding != HWPA_ENCODING_INT16 && encoding != HWPA_ENCODING_UINT8) return HWPA_ERR_INVALID_ARG;
    if (pool_id != a->poolId) return HWPA_ERR_BAD_POOL;
    if (first_frame != a->nextFrame) return HWPA_ERR_BAD_FRAME;
    const int L = a->lanes, F = a->frames;
    const size_t n = (size_t)L * F * LS;
    std::vector<int16_t> widened;
    const int16_t* tr = static_cast<const int16_t*>(traces);
    if (encoding == HWPA_ENCODING_UINT8) {
        widened.resize(n);
        const uint8_t* u = static_cast<const uint8_t*>(traces);
        for (size_t i = 0; i < n; ++i) widened[i] = (int16_t)u[i];
        tr = widened.data();
    }
    res->flags[0] = res->flags[1] = res->flags[2] = res->flags[3] = 0;
    int32_t rc = orc_baseliner_process(a->bl, tr, F, a->baselined.data(), a->stats.data());
    if (rc != HWPA_OK) return rc;
    a->nextFrame = first_frame + F;
    if (res->baselined) std::memcpy(res->baselined, a->baselined.data(), a->baselined.size() * 2);
    if (res->bl_stats) std::memcpy(res->bl_stats, a->stats.data(), sizeof(hwpa_baseliner_stats) * L);
    if (res->pulse_counts) std::memset(res->pulse_counts, 0, sizeof(uint32_t) * L * LS);
    if (res->pulse_overflows) std::memset(res->pulse_overflows, 0, sizeof(uint32_t) * L * LS);

    const int mode = a->cfg.model_estimation_mode;
    bool label = true, metrics = true;
    if (mode == HWPA_MODEL_INITIAL) {
        if (!a->modelInitialized && first_frame > (int32_t)a->baselinerLatency) {
            bool ran = false;
            if ((rc = OrcDmeAddBatch(a, first_frame, &ran)) != HWPA_OK) return rc;
            res->flags[1] = ran ? 1 : 0;
            a->modelInitialized = ran;
        }
        label = a->modelInitialized;
    } else if (mode == HWPA_MODEL_DYNAMIC) {
        if (first_frame < (int32_t)(a->baselinerLatency + a->dmeDelay)) return HWPA_OK;
        bool ran = false;
        if ((rc = OrcDmeAddBatch(a, first_frame, &ran)) != HWPA_OK) return rc;
        res->flags[1] = ran ? 1 : 0;
        if (ran) a->fullEstimation = true;
        metrics = a->fullEstimation;
    }
    if (res->models) std::memcpy(res->models, a->models.data(), sizeof(hwpa_lane_model) * L);

    const uint32_t mc = a->cfg.max_calls_per_zmw;
    std::vector<hwpa_pulse> pulses((size_t)L * LS * mc);
    std::vector<uint32_t> counts((size_t)L * LS, 0u), ovf((size_t)L * LS, 0u);
    std::vector<hwpa_stat_acc> pbl(L);
    std::memset(pbl.data(), 0, sizeof(hwpa_stat_acc) * L);
    if (label) {
        orc_labeler_process(a->lab, a->baselined.data(), F, a->models.data(), a->labels.data(), a->latent.data(),
                            a->score.data());
        res->flags[2] = 1;
        if (res->labels) std::memcpy(res->labels, a->labels.data(), a->labels.size() * 2);
        if (res->viterbi_score) std::memcpy(res->viterbi_score, a->score.data(), a->score.size() * 4);
        orc_pulse_process(a->pulse, a->labels.data(), a->latent.data(), a->baselined.data(), F,
                          first_frame - HWPA_LATENT_FRAMES, a->models.data(), pulses.data(), counts.data(),
                          ovf.data(), pbl.data());
    } else {
        // Empty labeler and pulse-detector metrics: Viterbi score 0 (FrameLabeler::EmptyMetrics),
        // no baseline frames.
        std::fill(a->score.begin(), a->score.end(), 0.0f);
    }
    if (!metrics) return HWPA_OK;
    res->flags[0] = label ? 1 : 0;
    if (label) {
        if (res->pulses) std::memcpy(res->pulses, pulses.data(), pulses.size() * sizeof(hwpa_pulse));
        if (res->pulse_counts) std::memcpy(res->pulse_counts, counts.data(), counts.size() * 4);
        if (res->pulse_overflows) std::memcpy(res->pulse_overflows, ovf.data(), ovf.size() * 4);
        if (res->pulse_bl_stats) std::memcpy(res->pulse_bl_stats, pbl.data(), sizeof(hwpa_stat_acc) * L);
    }
    // HF metrics of the pulse batch, which ends 24 latent frames before the chunk
    // (BatchAnalyzer.cpp:196-201, 275-280, 366-372).
    std::vector<hwpa_basecalling_metrics> met(L);
    res->flags[3] = orc_metrics_process(a->metrics, pulses.data(), counts.data(), mc, F,
                                        first_frame + F - HWPA_LATENT_FRAMES, a->stats.data(), a->score.data(),
                                        pbl.data(), a->models.data(), met.data());
    if (res->flags[3] && res->metrics) std::memcpy(res->metrics, met.data(), sizeof(hwpa_basecalling_metrics) * L);
    return HWPA_OK;
}

int32_t orc_analyzer_process(void* h, const int16_t* traces, uint32_t pool_id, int32_t first_frame,
                             orc_result* res)
{
    return orc_analyzer_process_encoded(h, traces, HWPA_ENCODING_INT16, pool_id, first_frame, res);
}

void orc_analyzer_set_em_form(void* h, int32_t form) { static_cast<OrcAnalyzer*>(h)->emForm = form; }

void orc_analyzer_set_activity_tree(void* h, const hwpa_activity_tree* t)
{
    orc_metrics_set_tree(static_cast<OrcAnalyzer*>(h)->metrics, t);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// BAZ packet encoding, restated from the reference's templates at run time
// (bazio/encoding/ObjectToBaz.h, FieldTransforms.h, FieldSerializers.h): a
// layout is a list of fields, each with a transform and a serialiser; fields
// are grouped as GenerateGroups does (ObjectToBaz.h:318-345) and every group is
// written as ceil(bits/8) little-endian bytes followed by its fields' overflow
// bytes in field order.
// ---------------------------------------------------------------------------

namespace {

enum { BZ_LABEL, BZ_START, BZ_WIDTH, BZ_ISBASE, BZ_PKMAX, BZ_PKMID, BZ_PKMEAN, BZ_PKVAR };
enum { BT_NOOP, BT_DELTA, BT_FLOATFIXED };
enum { BS_TRUNC, BS_SIMPLE, BS_COMPACT };

struct BzField {
    int name;
    bool sgn;
    int trans;
    uint32_t scale;
    int tbytes;
    int ser;
    int bits;
    int obytes;
};

struct BzLayout {
    std::vector<BzField> f;
    std::vector<int> groupStart;  // first field of each group; groupStart.back() == f.size()
    std::vector<int> groupBytes;
};

static BzLayout MakeLayout(std::vector<BzField> fields)
{
    BzLayout L;
    L.f = std::move(fields);
    int cur = 0;
    L.groupStart.push_back(0);
    for (size_t k = 0; k < L.f.size(); ++k) {
        const int next = cur + L.f[k].bits;
        const bool evenBytes = cur % 8 == 0 && cur > 0;
        if (evenBytes || next > 64) {
            L.groupBytes.push_back((cur + 7) / 8);
            L.groupStart.push_back((int)k);
            cur = L.f[k].bits;
        } else {
            cur = next;
        }
    }
    L.groupBytes.push_back((cur + 7) / 8);
    L.groupStart.push_back((int)L.f.size());
    return L;
}

// Layouts: 0 ProductionPulses, 1 InternalPulses (basecaller/dataTypes/PulseGroups.h:41-94);
// 2-4 the lossy-truncate, simple-overflow and compact-overflow layouts of the reference's
// PulseToBaz tests (bazio/encoding/test/TestPulseToBaz.cu:110-300).
static const BzLayout& Layout(int id)
{
    static const BzLayout ls[5] = {
        MakeLayout({{BZ_LABEL, false, BT_NOOP, 0, 0, BS_TRUNC, 2, 0},
                    {BZ_WIDTH, false, BT_NOOP, 0, 0, BS_COMPACT, 7, 0},
                    {BZ_START, false, BT_DELTA, 0, 0, BS_COMPACT, 7, 0}}),
        MakeLayout({{BZ_LABEL, false, BT_NOOP, 0, 0, BS_TRUNC, 2, 0},
                    {BZ_WIDTH, false, BT_NOOP, 0, 0, BS_COMPACT, 7, 0},
                    {BZ_START, false, BT_DELTA, 0, 0, BS_COMPACT, 7, 0},
                    {BZ_PKMAX, true, BT_FLOATFIXED, 10, 2, BS_SIMPLE, 8, 2},
                    {BZ_PKMID, true, BT_FLOATFIXED, 10, 2, BS_SIMPLE, 8, 2},
                    {BZ_PKMEAN, true, BT_FLOATFIXED, 10, 2, BS_SIMPLE, 8, 2},
                    {BZ_PKVAR, false, BT_FLOATFIXED, 10, 2, BS_SIMPLE, 7, 2},
                    {BZ_ISBASE, false, BT_NOOP, 0, 0, BS_TRUNC, 1, 0}}),
        MakeLayout({{BZ_LABEL, false, BT_NOOP, 0, 0, BS_TRUNC, 2, 0},
                    {BZ_WIDTH, false, BT_NOOP, 0, 0, BS_TRUNC, 7, 0},
                    {BZ_START, false, BT_DELTA, 0, 0, BS_TRUNC, 7, 0}}),
        MakeLayout({{BZ_LABEL, false, BT_NOOP, 0, 0, BS_TRUNC, 2, 0},
                    {BZ_WIDTH, false, BT_NOOP, 0, 0, BS_SIMPLE, 7, 4},
                    {BZ_START, false, BT_DELTA, 0, 0, BS_SIMPLE, 7, 4}}),
        MakeLayout({{BZ_LABEL, false, BT_NOOP, 0, 0, BS_TRUNC, 2, 0},
                    {BZ_WIDTH, false, BT_NOOP, 0, 0, BS_COMPACT, 7, 0},
                    {BZ_START, false, BT_DELTA, 0, 0, BS_COMPACT, 7, 0}}),
    };
    return ls[id];
}

// FloatFixedCodec special codes and offset (FieldTransforms.h:201-207).
constexpr uint64_t kFfNan = 0, kFfInf = 1, kFfNinf = 2, kFfOffset = 3;

static uint64_t FloatFixedApply(float f, bool sgn, uint32_t scale, int bytes)
{
    float scaled = std::round(f * (float)scale);
    if (sgn) {
        const int64_t mx = (int64_t)((1ull << (8 * bytes - 1)) - 1 - kFfOffset);
        if ((float)(-mx) <= scaled && (float)mx >= scaled) {
            if (scaled >= 0) scaled += (float)kFfOffset;
            return (uint64_t)(int64_t)scaled;
        }
    } else {
        const uint64_t mx = (bytes >= 8 ? ~0ull : (1ull << (8 * bytes)) - 1) - kFfOffset;
        if (0 <= scaled && (float)mx >= scaled) {
            scaled += (float)kFfOffset;
            return (uint64_t)scaled;
        }
    }
    if (std::isnan(f)) return kFfNan;
    return scaled < 0 ? kFfNinf : kFfInf;
}

static float FloatFixedRevert(uint64_t v, bool sgn, uint32_t scale)
{
    if (v == kFfNan) return NAN;
    if (v == kFfInf) return INFINITY;
    if (v == kFfNinf) return -INFINITY;
    if (sgn) {
        int64_t s = (int64_t)v;
        if (s >= 0) s -= (int64_t)kFfOffset;
        return (float)s / (float)scale;
    }
    return (float)(v - kFfOffset) / (float)scale;
}

// Overflow bytes a serialiser needs for v (FieldSerializers.h:57-263).
static int OverflowBytes(const BzField& f, uint64_t v)
{
    if (f.ser == BS_TRUNC) return 0;
    if (f.ser == BS_SIMPLE) {
        uint64_t t = v;
        if (f.sgn) t += (1ull << (f.bits - 1)) - 1;
        return t >= (1ull << f.bits) - 1 ? f.obytes : 0;
    }
    if (f.sgn) {
        const int discard = __builtin_clrsbl((long)v);
        const int keep = 64 - discard - (f.bits - 1);
        return (keep + 6) / 7;
    }
    uint64_t t = v >> (f.bits - 1);
    int n = 0;
    while (t) {
        ++n;
        t >>= 7;
    }
    return n;
}

static uint64_t ToBinary(const BzField& f, uint64_t v, uint8_t*& p, bool write)
{
    if (f.ser == BS_TRUNC) return v;
    if (f.ser == BS_SIMPLE) {
        if (OverflowBytes(f, v)) {
            if (write) std::memcpy(p, &v, (size_t)f.obytes);
            p += f.obytes;
            return (1ull << f.bits) - 1;
        }
        return f.sgn ? v + ((1ull << (f.bits - 1)) - 1) : v;
    }
    const uint64_t contMask = 1ull << (f.bits - 1);
    int n = OverflowBytes(f, v);
    uint64_t main = v & (contMask - 1);
    v >>= (f.bits - 1);
    if (n) main |= contMask;
    while (n) {
        uint8_t b = (uint8_t)(v & 0x7F);
        v >>= 7;
        --n;
        if (n) b |= 0x80;
        if (write) *p = b;
        ++p;
    }
    return main;
}

static uint64_t FromBinary(const BzField& f, uint64_t v, const uint8_t*& p)
{
    if (f.ser == BS_TRUNC) {
        if (f.sgn && (v & (1ull << (f.bits - 1)))) v |= ~0ull << f.bits;
        return v;
    }
    if (f.ser == BS_SIMPLE) {
        if (v == (1ull << f.bits) - 1) {
            uint64_t r = 0;
            std::memcpy(&r, p, (size_t)f.obytes);
            p += f.obytes;
            const int signBit = 8 * f.obytes;
            if (f.sgn && (r & (1ull << (signBit - 1)))) r |= ~0ull << (signBit - 1);
            return r;
        }
        return f.sgn ? v - ((1ull << (f.bits - 1)) - 1) : v;
    }
    const uint64_t contMask = 1ull << (f.bits - 1);
    uint64_t r = v & (contMask - 1);
    bool cont = (v & contMask) != 0;
    int cur = f.bits - 1;
    while (cont) {
        const uint64_t b = *p++;
        cont = (b & 0x80) != 0;
        r |= (b & 0x7F) << cur;
        cur += 7;
    }
    if (f.sgn && (r & (1ull << (cur - 1)))) r |= ~0ull << cur;
    return r;
}

struct BzValues {
    uint64_t label, start, width, isbase;
    float pkmax, pkmid, pkmean, pkvar;
};

static uint64_t Apply(const BzField& f, const BzValues& v, uint64_t& last)
{
    uint64_t x = 0;
    float fx = 0.0f;
    switch (f.name) {
    case BZ_LABEL: x = v.label; break;
    case BZ_START: x = v.start; break;
    case BZ_WIDTH: x = v.width; break;
    case BZ_ISBASE: x = v.isbase; break;
    case BZ_PKMAX: fx = v.pkmax; break;
    case BZ_PKMID: fx = v.pkmid; break;
    case BZ_PKMEAN: fx = v.pkmean; break;
    default: fx = v.pkvar; break;
    }
    if (f.trans == BT_FLOATFIXED) return FloatFixedApply(fx, f.sgn, f.scale, f.tbytes);
    if (f.trans == BT_DELTA) {
        const uint64_t d = x - last;
        last = x;
        return d;
    }
    return x;
}

// ObjectEncoder::Serialize (ObjectToBaz.h:259-265) of one object; write = false only counts.
static size_t EncodeOne(const BzLayout& L, const BzValues& v, uint64_t& last, uint8_t* out, bool write)
{
    uint8_t* p = out;
    for (size_t g = 0; g + 1 < L.groupStart.size(); ++g) {
        const int nb = L.groupBytes[g];
        uint8_t* p2 = p + nb;
        uint64_t data = 0;
        int pos = 0;
        for (int k = L.groupStart[g]; k < L.groupStart[g + 1]; ++k) {
            const BzField& f = L.f[k];
            const uint64_t main = ToBinary(f, Apply(f, v, last), p2, write);
            data |= (main & ((1ull << f.bits) - 1)) << pos;
            pos += f.bits;
        }
        if (write) std::memcpy(p, &data, (size_t)nb);
        p = p2;
    }
    return (size_t)(p - out);
}

static size_t DecodeOne(const BzLayout& L, const uint8_t* in, uint64_t& last, BzValues& v)
{
    const uint8_t* p = in;
    for (size_t g = 0; g + 1 < L.groupStart.size(); ++g) {
        const int nb = L.groupBytes[g];
        uint64_t data = 0;
        std::memcpy(&data, p, (size_t)nb);
        p += nb;
        for (int k = L.groupStart[g]; k < L.groupStart[g + 1]; ++k) {
            const BzField& f = L.f[k];
            const uint64_t raw = data & ((1ull << f.bits) - 1);
            data >>= f.bits;
            uint64_t x = FromBinary(f, raw, p);
            if (f.trans == BT_DELTA) {
                last += x;
                x = last;
            }
            switch (f.name) {
            case BZ_LABEL: v.label = x; break;
            case BZ_START: v.start = x; break;
            case BZ_WIDTH: v.width = x; break;
            case BZ_ISBASE: v.isbase = x; break;
            default: {
                const float fv = f.trans == BT_FLOATFIXED ? FloatFixedRevert(x, f.sgn, f.scale) : (float)x;
                if (f.name == BZ_PKMAX) v.pkmax = fv;
                else if (f.name == BZ_PKMID) v.pkmid = fv;
                else if (f.name == BZ_PKMEAN) v.pkmean = fv;
                else v.pkvar = fv;
            }
            }
        }
    }
    return (size_t)(p - in);
}

// FieldAccessor<Pulse, PacketFieldName>::Get (basecaller/dataTypes/PulseFieldAccessors.h:40-70)
// on the fixed-point pulse record (dataTypes/Pulse.h:104-126).
static BzValues PulseValues(const hwpa_pulse& p)
{
    BzValues v;
    v.label = p.label;
    v.start = p.start;
    v.width = p.width;
    v.isbase = p.is_reject ? 0 : 1;
    v.pkmax = (float)p.max_signal * 0.1f;
    v.pkmean = (float)p.mean_signal * 0.1f;
    v.pkmid = p.width < 3 ? NAN : (float)p.mid_signal * 0.1f;
    v.pkvar = p.width < 3 ? NAN : p.signal_m2;
    return v;
}

}  // namespace

extern "C" {

int64_t orc_baz_encode_fields(int32_t layout, int32_t n, const uint8_t* label, const uint32_t* start,
                              const uint16_t* width, const uint8_t* isbase, const float* pkmax, const float* pkmid,
                              const float* pkmean, const float* pkvar, uint64_t* last, uint8_t* out)
{
    const BzLayout& L = Layout(layout);
    size_t off = 0;
    for (int32_t i = 0; i < n; ++i) {
        BzValues v{label[i], start[i], width[i], isbase ? isbase[i] : 1u,
                   pkmax ? pkmax[i] : 0.0f, pkmid ? pkmid[i] : 0.0f, pkmean ? pkmean[i] : 0.0f, pkvar ? pkvar[i] : 0.0f};
        off += EncodeOne(L, v, *last, out ? out + off : nullptr, out != nullptr);
    }
    return (int64_t)off;
}

int64_t orc_baz_decode_fields(int32_t layout, int32_t n, const uint8_t* in, uint64_t* last, uint8_t* label,
                              uint32_t* start, uint16_t* width, uint8_t* isbase, float* pkmax, float* pkmid,
                              float* pkmean, float* pkvar)
{
    const BzLayout& L = Layout(layout);
    size_t off = 0;
    for (int32_t i = 0; i < n; ++i) {
        BzValues v{0, 0, 0, 1, 0.0f, 0.0f, 0.0f, 0.0f};
        off += DecodeOne(L, in + off, *last, v);
        label[i] = (uint8_t)v.label;
        start[i] = (uint32_t)v.start;
        width[i] = (uint16_t)v.width;
        if (isbase) isbase[i] = (uint8_t)v.isbase;
        if (pkmax) pkmax[i] = v.pkmax;
        if (pkmid) pkmid[i] = v.pkmid;
        if (pkmean) pkmean[i] = v.pkmean;
        if (pkvar) pkvar[i] = v.pkvar;
    }
    return (int64_t)off;
}

// Group bit widths of a layout (GroupParams::numBits / totalBits); returns the group count.
int32_t orc_baz_layout_groups(int32_t layout, int32_t* group_bits, int32_t* group_fields, int32_t cap)
{
    const BzLayout& L = Layout(layout);
    const int ng = (int)L.groupBytes.size();
    for (int g = 0; g < ng && g < cap; ++g) {
        int bits = 0;
        for (int k = L.groupStart[g]; k < L.groupStart[g + 1]; ++k) bits += L.f[k].bits;
        group_bits[g] = bits;
        group_fields[g] = L.groupStart[g + 1] - L.groupStart[g];
    }
    return ng;
}

// One serialiser on its own: ToBinary, the main portion masked to `bits`, FromBinary
// (the round trips of bazio/encoding/test/TestFieldSerializers.cpp).
uint64_t orc_baz_serial_roundtrip(int32_t ser, int32_t bits, int32_t obytes, int32_t sgn, uint64_t v,
                                  int32_t* overflow_bytes)
{
    const BzField f{BZ_START, sgn != 0, BT_NOOP, 0, 0, ser, bits, obytes};
    uint8_t buf[16] = {0};
    uint8_t* p = buf;
    if (overflow_bytes) *overflow_bytes = OverflowBytes(f, v);
    uint64_t main = ToBinary(f, v, p, true) & (bits >= 64 ? ~0ull : (1ull << bits) - 1);
    const uint8_t* q = buf;
    return FromBinary(f, main, q);
}

uint64_t orc_baz_floatfixed_apply(float f, int32_t sgn, uint32_t scale, int32_t bytes)
{
    return FloatFixedApply(f, sgn != 0, scale, bytes);
}

float orc_baz_floatfixed_revert(uint64_t v, int32_t sgn, uint32_t scale)
{
    return FloatFixedRevert(v, sgn != 0, scale);
}

// A chunk of pulses [zmws][max_calls] into one stream ordered by ZMW, as hwpa_baz_encode
// (production: bases only, PrelimHQFilter.cpp:148-160).  Returns the total; writes nothing
// and keeps the delta state when it exceeds cap.
int64_t orc_baz_encode_batch(int32_t internal, const hwpa_pulse* pulses, const uint32_t* counts, uint32_t max_calls,
                             int32_t zmws, uint64_t* last, uint8_t* out, uint64_t cap, uint64_t* offsets,
                             uint32_t* events)
{
    const BzLayout& L = Layout(internal ? 1 : 0);
    uint64_t total = 0;
    for (int32_t z = 0; z < zmws; ++z) {
        offsets[z] = total;
        uint64_t prev = last[z];
        uint32_t ev = 0;
        const uint32_t n = std::min(counts[z], max_calls);
        for (uint32_t i = 0; i < n; ++i) {
            const hwpa_pulse& p = pulses[(size_t)z * max_calls + i];
            if (!internal && p.is_reject) continue;
            total += EncodeOne(L, PulseValues(p), prev, nullptr, false);
            ++ev;
        }
        events[z] = ev;
    }
    offsets[zmws] = total;
    if (total > cap) return (int64_t)total;
    for (int32_t z = 0; z < zmws; ++z) {
        uint8_t* o = out + offsets[z];
        const uint32_t n = std::min(counts[z], max_calls);
        for (uint32_t i = 0; i < n; ++i) {
            const hwpa_pulse& p = pulses[(size_t)z * max_calls + i];
            if (!internal && p.is_reject) continue;
            o += EncodeOne(L, PulseValues(p), last[z], o, true);
        }
    }
    return (int64_t)total;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Hooks used by the known-answer tests
// ---------------------------------------------------------------------------

extern "C" {

// Element-wise evaluation of the shared math routines of include/hwpa_math.h.
// kind: 0 exp, 1 log, 2 log2, 3 exp2, 4 pow(a,b), 5 f->half->f, 6 round_i16, 7 floor_i32,
// 9 / 10 table-driven log / reciprocal of a normal float, 11 / 12 Cephes exp / log
// (the reference's LaneArray exp / log), 13 / 14 / 15 libm exp2f / log2f / powf,
// 16 / 17 / 18 the kernels' double-rounded exp2 / log2 / pow
void orc_math(int32_t kind, const float* a, const float* b, float* out, int32_t n)
{
    for (int32_t i = 0; i < n; ++i) {
        switch (kind) {
        case 11: out[i] = hw_cephes_expf(a[i]); break;
        case 12: out[i] = hw_cephes_logf(a[i]); break;
        case 13: out[i] = std::exp2(a[i]); break;
        case 14: out[i] = std::log2(a[i]); break;
        case 15: out[i] = std::pow(a[i], b[i]); break;
        case 16: out[i] = hw_dexp2f(a[i]); break;
        case 17: out[i] = hw_dlog2f(a[i]); break;
        case 18: out[i] = hw_dpowf(a[i], b[i]); break;
        case 0: out[i] = hw_expf(a[i]); break;
        case 1: out[i] = hw_logf(a[i]); break;
        case 2: out[i] = hw_log2f(a[i]); break;
        case 3: out[i] = hw_exp2f(a[i]); break;
        case 4: out[i] = hw_powf(a[i], b[i]); break;
        case 5: out[i] = hw_h2f(hw_f2h(a[i])); break;
        case 6: out[i] = (float)hw_round_i16(a[i]); break;
        case 9:
        case 10: {
            float lg, rcp;
            hw_log_rcp_tab(a[i], LogTab().rc, LogTab().lc, &lg, &rcp);
            out[i] = kind == 9 ? lg : rcp;
            break;
        }
        default: out[i] = (float)hw_floor_i32(a[i]); break;
        }
    }
}

// Runs one fresh extremum filter (ExtremumFilterHgw + BlockFilterStage) over
// `chunks` consecutive blocks of `len` lane-frames; outputs are compacted per
// block into out[chunk][k][64] with k < ceil(len / stride).
void orc_hgw(int32_t width, int32_t stride, int32_t is_max, int32_t chunks, int32_t len, const int16_t* x,
             int16_t* out)
{
    HgwFilter f;
    f.Init(width, stride, 1, is_max != 0);
    std::vector<int16_t> buf((size_t)len * LS);
    const int nout = (len + stride - 1) / stride;
    for (int c = 0; c < chunks; ++c) {
        std::memcpy(buf.data(), x + (size_t)c * len * LS, buf.size() * 2);
        f.Run(buf.data(), len);
        std::memcpy(out + (size_t)c * nout * LS, buf.data(), (size_t)nout * LS * 2);
    }
}

// The Viterbi recursion of FrameLabelerHost.cpp:213-264 over given per-frame
// state scores: scores[f][13], roi[f], initial log-likelihoods ll0[13].
// Writes the traceback-decoded best path ending in argmax of the final
// log-likelihood (later state wins ties) and the final log-likelihoods.
// ---- single-function hooks for the known-answer tests -----------------------------
// Models as 16 floats: baseline {weight, mean, var}, analogs[4] {weight, mean, var}, confidence.
static_assert(sizeof(ModelZ) == 16 * sizeof(float), "ModelZ is 16 packed floats");

// DetectionModelHost::Update(other, fraction) (fraction >= 0) or Update(other) (fraction < 0).
void orc_dme_update(int32_t form, const hwpa_config* cfg, const float* t16, const float* o16, float fraction,
                    float* out16)
{
    const DmeCfg d = MakeDmeCfg(*cfg);
    ModelZ t, o;
    std::memcpy(&t, t16, sizeof t);
    std::memcpy(&o, o16, sizeof o);
    if (form == ORC_FORM_MIRROR) {
        if (fraction < 0.0f) UpdateConf<MirrorForm>(d, t, o);
        else UpdateFraction<MirrorForm>(d, t, o, fraction);
    } else {
        if (fraction < 0.0f) UpdateConf<ReferenceForm>(d, t, o);
        else UpdateFraction<ReferenceForm>(d, t, o, fraction);
    }
    std::memcpy(out16, &t, sizeof t);
}

// DmeEmHost::ComputeConfidence: the 8 factors for one ZMW.
void orc_dme_confidence(int32_t form, const hwpa_config* cfg, int32_t converged, const float* ref16,
                        const float* est16, float* cf8)
{
    const DmeCfg d = MakeDmeCfg(*cfg);
    ModelZ r, e;
    std::memcpy(&r, ref16, sizeof r);
    std::memcpy(&e, est16, sizeof e);
    if (form == ORC_FORM_MIRROR) ComputeConfidence<MirrorForm>(d, converged != 0, r, e, cf8);
    else ComputeConfidence<ReferenceForm>(d, converged != 0, r, e, cf8);
}

// EvolveModel's confidence multiplier between two frame intervals [lo, hi).
float orc_dme_evolve(int32_t form, const hwpa_config* cfg, const int32_t model_fi[2], const int32_t est_fi[2])
{
    const DmeCfg d = MakeDmeCfg(*cfg);
    return form == ORC_FORM_MIRROR ? EvolveFactor<MirrorForm>(d, model_fi, est_fi)
                                   : EvolveFactor<ReferenceForm>(d, model_fi, est_fi);
}

// UHistogramSimd::Fractile / CumulativeCount per ZMW of one lane histogram.
void orc_hist_fractile(const hwpa_lane_hist* lh, const float* frac, float* out)
{
    HistZ h;
    for (int z = 0; z < LS; ++z) {
        MakeHistZ(*lh, z, h);
        out[z] = h.Fractile(frac[z]);
    }
}
void orc_hist_cumulative(const hwpa_lane_hist* lh, const float* x, float* out)
{
    HistZ h;
    for (int z = 0; z < LS; ++z) {
        MakeHistZ(*lh, z, h);
        out[z] = h.CumulativeCount(x[z]);
    }
}

// DmeEmHost::PrelimScaleFactor per ZMW: models16 [64][16].
void orc_prelim_scale_factor(const hwpa_config* cfg, const hwpa_lane_hist* lh, const float* models16, float* out)
{
    const DmeCfg d = MakeDmeCfg(*cfg);
    HistZ h;
    for (int z = 0; z < LS; ++z) {
        MakeHistZ(*lh, z, h);
        ModelZ m;
        std::memcpy(&m, models16 + 16 * z, sizeof m);
        out[z] = PrelimScaleFactor(d, m, h);
    }
}

// BlockStateSubframeScorer::StateScores of ZMW z for n data values: out [n][13].
void orc_scorer_scores(const hwpa_lane_model* m, int32_t z, const float* x, int32_t n, float* out)
{
    Scorer sc;
    sc.Setup(*m, z);
    for (int32_t i = 0; i < n; ++i)
        for (int s = 0; s < NS; ++s) out[(size_t)i * NS + s] = sc.Score(x[i], s);
}

// The labeler's ROI of one ZMW: seq [nroi + 2] sigma-normalised frames after `before`.
void orc_roi(float upper, float lower, float before, uint16_t boundary, const float* seq, int32_t nroi, uint8_t* out)
{
    std::vector<uint16_t> r((size_t)nroi);
    RoiZ(upper, lower, before, boundary, [&](int k) { return seq[k]; }, nroi, r.data(), 1);
    for (int32_t k = 0; k < nroi; ++k) out[k] = (uint8_t)r[k];
}

void orc_viterbi_core(const hwpa_config* cfg, int32_t frames, const float* scores, const uint8_t* roi,
                      const float* ll0, int16_t* path, float* llOut)
{
    hwpa_transition T;
    BuildTransition(*cfg, T);
    float ll[NS];
    for (int s = 0; s < NS; ++s) ll[s] = ll0[s];
    std::vector<uint8_t> tb((size_t)frames * NS);
    for (int f = 0; f < frames; ++f) {
        float acc[NS];
        for (int r = 0; r < NS; ++r) {
            const float sco = scores[f * NS + r];
            const int fc = T.first_col[r];
            float mx = sco + T.logp[r][fc] + ll[fc];
            int mi = fc;
            for (int p = T.seg_first[r]; p < T.seg_last[r]; ++p) {
                const float v = sco + T.logp[r][p] + ll[p];
                if (v >= mx) {
                    mx = v;
                    mi = p;
                }
            }
            acc[r] = mx - ((r == 0 || roi[f]) ? 0.0f : INFINITY);
            tb[(size_t)f * NS + r] = (uint8_t)mi;
        }
        for (int r = 0; r < NS; ++r) ll[r] = acc[r];
    }
    int st = 0;
    for (int s = 1; s < NS; ++s)
        if (!(ll[st] > ll[s])) st = s;
    for (int f = frames - 1; f >= 0; --f) {
        path[f] = (int16_t)st;
        st = tb[(size_t)f * NS + st];
    }
    for (int s = 0; s < NS; ++s) llOut[s] = ll[s];
}

}  // extern "C"
