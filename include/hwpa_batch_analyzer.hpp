/*
 * hwpa_batch_analyzer.hpp -- the reference's per-pool analyzer shape over the C-ABI.
 *
 * hwpa::BatchAnalyzer mirrors analyzer/BatchAnalyzer.h:51-106: constructed once per
 * pool, called once per chunk with a trace batch, returning a BatchResult
 * (dataTypes/BatchResult.h) that owns a PulseBatch -- the 20-byte Pulse records of
 * every ZMW with their counts and overflow flags (dataTypes/BatchVectors.h) -- and
 * the chunk's BasecallingMetrics when a metrics block ended (BatchResult::metrics,
 * analyzer/BatchAnalyzer.cpp:366-372).  Failures become hwpa::Error exceptions that
 * carry the library's status code and message, as the reference raises
 * PBException.  Header-only host C++17; it links against libhwpa.so and needs no
 * HIP headers (results are copied with the library's own device helpers).
 */
#ifndef HWPA_BATCH_ANALYZER_HPP
#define HWPA_BATCH_ANALYZER_HPP

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "hwpa.h"

namespace hwpa {

class Error : public std::runtime_error {
public:
    Error(int32_t code, const std::string& msg) : std::runtime_error(msg), code_(code) {}
    int32_t Code() const { return code_; }

private:
    int32_t code_;
};

inline void Check(int32_t rc)
{
    if (rc != HWPA_OK) throw Error(rc, hwpa_last_error());
}

/* One chunk of traces: TraceBatchVariant's data plus the BatchMetadata the analyzer
 * checks (pool id, first frame); samples [lanes][frames][64] of `encoding`, in host
 * memory or already in device memory. */
struct TraceBatch {
    uint32_t poolId = 0;
    int32_t firstFrame = 0;
    int32_t encoding = HWPA_ENCODING_INT16;
    const void* data = nullptr;
    bool onDevice = false;
};

/* PulseBatch: per-ZMW pulse slots [zmws][maxCallsPerZmw], their counts and overflows. */
struct PulseBatch {
    uint32_t poolId = 0;
    int32_t firstFrame = 0, lastFrame = 0;   // frames the pulses cover (24 latent frames earlier)
    uint32_t zmws = 0, maxCallsPerZmw = 0;
    bool empty = true;                       // start-up window: no pulses produced
    std::vector<hwpa_pulse> pulses;
    std::vector<uint32_t> lengths, overflows;
    const hwpa_pulse* ZmwPulses(uint32_t z) const { return pulses.data() + (size_t)z * maxCallsPerZmw; }
};

struct BatchResult {
    PulseBatch pulses;
    std::unique_ptr<std::vector<hwpa_basecalling_metrics>> metrics;   // null unless a block ended
};

class BatchAnalyzer {
public:
    BatchAnalyzer(uint32_t poolId, uint32_t lanes, uint32_t framesPerChunk, const hwpa_config& cfg)
        : poolId_(poolId), lanes_(lanes), frames_(framesPerChunk)
    {
        Check(hwpa_analyzer_create(poolId, lanes, framesPerChunk, &cfg, &h_));
    }
    ~BatchAnalyzer() { hwpa_analyzer_destroy(h_); }
    BatchAnalyzer(const BatchAnalyzer&) = delete;
    BatchAnalyzer& operator=(const BatchAnalyzer&) = delete;

    uint32_t PoolId() const { return poolId_; }
    hwpa_analyzer* Handle() const { return h_; }

    /* BatchAnalyzer::operator() (analyzer/BatchAnalyzer.cpp:149-170). */
    BatchResult operator()(const TraceBatch& tb)
    {
        Check(hwpa_analyzer_process_encoded(h_, tb.data, tb.encoding, tb.onDevice ? 1 : 0, tb.poolId,
                                            tb.firstFrame));
        Check(hwpa_analyzer_sync(h_));
        int32_t flags[4];
        Check(hwpa_result_flags(h_, flags));
        BatchResult r;
        PulseBatch& pb = r.pulses;
        pb.poolId = tb.poolId;
        pb.firstFrame = tb.firstFrame - HWPA_LATENT_FRAMES;
        pb.lastFrame = tb.firstFrame + (int32_t)frames_ - HWPA_LATENT_FRAMES;
        pb.zmws = lanes_ * HWPA_LANE_SIZE;
        pb.empty = flags[0] == 0;
        const hwpa_pulse* dPulses = nullptr;
        const uint32_t* dCounts = nullptr;
        const uint32_t* dOverflows = nullptr;
        Check(hwpa_result_pulses(h_, &dPulses, &dCounts, &dOverflows, &pb.maxCallsPerZmw));
        pb.lengths.assign(pb.zmws, 0u);
        pb.overflows.assign(pb.zmws, 0u);
        pb.pulses.resize((size_t)pb.zmws * pb.maxCallsPerZmw);
        if (!pb.empty) {
            Check(hwpa_memcpy_dtoh(pb.lengths.data(), dCounts, sizeof(uint32_t) * pb.zmws));
            Check(hwpa_memcpy_dtoh(pb.overflows.data(), dOverflows, sizeof(uint32_t) * pb.zmws));
            Check(hwpa_memcpy_dtoh(pb.pulses.data(), dPulses, sizeof(hwpa_pulse) * pb.pulses.size()));
        }
        if (flags[3]) {
            const hwpa_basecalling_metrics* dMetrics = nullptr;
            Check(hwpa_result_metrics(h_, &dMetrics));
            r.metrics.reset(new std::vector<hwpa_basecalling_metrics>(lanes_));
            Check(hwpa_memcpy_dtoh(r.metrics->data(), dMetrics, sizeof(hwpa_basecalling_metrics) * lanes_));
        }
        return r;
    }

private:
    uint32_t poolId_, lanes_, frames_;
    hwpa_analyzer* h_ = nullptr;
};

}  // namespace hwpa

#endif /* HWPA_BATCH_ANALYZER_HPP */
