// Drives hwpa::BatchAnalyzer (include/hwpa_batch_analyzer.hpp) the way the reference's
// BasecallerBody drives its BatchAnalyzer (appModules/Basecaller.h:277-281) and checks
// every BatchResult against the C-ABI analyzer run on the same chunks: the PulseBatch
// (records, counts, overflows) byte for byte, the forwarded metrics whenever a block
// ends, and the wrong-pool / out-of-order errors as exceptions.  Exit status 0 = pass.
#include <cstdio>
#include <cstring>
#include <vector>

#include "hwpa_batch_analyzer.hpp"

#define REQUIRE(c)                                                            \
    do {                                                                      \
        if (!(c)) {                                                           \
            std::fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #c); \
            return 1;                                                         \
        }                                                                     \
    } while (0)

int main()
{
    const uint32_t lanes = 4, frames = 512, chunks = 12, pool = 5;
    hwpa_config cfg;
    hwpa::Check(hwpa_config_default(&cfg));
    // an estimate every chunk, short start-up; metrics blocks of 1024 frames
    hwpa::Check(hwpa_config_load_json(&cfg, "{\"baselinerConfig\":{\"SigmaEmaScaleStrides\":0},"
                                            "\"dmeConfig\":{\"MinFramesForEstimate\":512,"
                                            "\"NumFramesPreAccumStats\":512}}"));
    cfg.frames_per_hf_metric_block = 1024;
    hwpa_tracegen_params gen;
    hwpa::Check(hwpa_tracegen_default(&gen));
    void* dState = nullptr;
    void* dTraces = nullptr;
    const size_t chunkBytes = (size_t)lanes * frames * HWPA_LANE_SIZE * 2;
    hwpa::Check(hwpa_dev_alloc(&dState, sizeof(hwpa_tracegen_state) * lanes));
    hwpa::Check(hwpa_dev_alloc(&dTraces, chunkBytes));

    hwpa::BatchAnalyzer wrapped(pool, lanes, frames, cfg);
    hwpa_analyzer* direct = nullptr;
    hwpa::Check(hwpa_analyzer_create(pool, lanes, frames, &cfg, &direct));
    int produced = 0, metricBlocks = 0;
    for (uint32_t c = 0; c < chunks; ++c) {
        hwpa::Check(hwpa_tracegen_run(&gen, lanes, 0, frames, c * frames, (hwpa_tracegen_state*)dState, c == 0,
                                      (int16_t*)dTraces, nullptr));
        hwpa::Check(hwpa_device_sync());
        hwpa::TraceBatch tb;
        tb.poolId = pool;
        tb.firstFrame = (int32_t)(c * frames);
        tb.data = dTraces;
        tb.onDevice = true;
        const hwpa::BatchResult r = wrapped(tb);
        hwpa::Check(hwpa_analyzer_process(direct, (const int16_t*)dTraces, 1, pool, tb.firstFrame));
        hwpa::Check(hwpa_analyzer_sync(direct));
        int32_t flags[4];
        hwpa::Check(hwpa_result_flags(direct, flags));
        REQUIRE(r.pulses.empty == (flags[0] == 0));
        REQUIRE((r.metrics != nullptr) == (flags[3] != 0));
        REQUIRE(r.pulses.poolId == pool && r.pulses.zmws == lanes * HWPA_LANE_SIZE);
        if (!r.pulses.empty) {
            const hwpa_pulse* dp;
            const uint32_t *dc, *dov;
            uint32_t mc;
            hwpa::Check(hwpa_result_pulses(direct, &dp, &dc, &dov, &mc));
            REQUIRE(mc == r.pulses.maxCallsPerZmw);
            std::vector<uint32_t> cnt(r.pulses.zmws), ovf(r.pulses.zmws);
            std::vector<hwpa_pulse> p((size_t)r.pulses.zmws * mc);
            hwpa::Check(hwpa_memcpy_dtoh(cnt.data(), dc, 4 * cnt.size()));
            hwpa::Check(hwpa_memcpy_dtoh(ovf.data(), dov, 4 * ovf.size()));
            hwpa::Check(hwpa_memcpy_dtoh(p.data(), dp, sizeof(hwpa_pulse) * p.size()));
            REQUIRE(cnt == r.pulses.lengths && ovf == r.pulses.overflows);
            for (uint32_t z = 0; z < r.pulses.zmws; ++z) {
                const uint32_t k = cnt[z] < mc ? cnt[z] : mc;
                REQUIRE(std::memcmp(r.pulses.ZmwPulses(z), &p[(size_t)z * mc], sizeof(hwpa_pulse) * k) == 0);
                produced += (int)k;
            }
        }
        if (r.metrics) {
            const hwpa_basecalling_metrics* dm;
            hwpa::Check(hwpa_result_metrics(direct, &dm));
            std::vector<hwpa_basecalling_metrics> m(lanes);
            hwpa::Check(hwpa_memcpy_dtoh(m.data(), dm, sizeof(hwpa_basecalling_metrics) * lanes));
            REQUIRE(std::memcmp(m.data(), r.metrics->data(), sizeof(hwpa_basecalling_metrics) * lanes) == 0);
            ++metricBlocks;
        }
    }
    // BatchAnalyzer.cpp:151-152: wrong pool, then a chunk out of order.
    hwpa::TraceBatch bad;
    bad.poolId = pool + 1;
    bad.firstFrame = (int32_t)(chunks * frames);
    bad.data = dTraces;
    bad.onDevice = true;
    int32_t code = HWPA_OK;
    try {
        wrapped(bad);
    } catch (const hwpa::Error& e) {
        code = e.Code();
    }
    REQUIRE(code == HWPA_ERR_BAD_POOL);
    bad.poolId = pool;
    bad.firstFrame = 0;
    code = HWPA_OK;
    try {
        wrapped(bad);
    } catch (const hwpa::Error& e) {
        code = e.Code();
    }
    REQUIRE(code == HWPA_ERR_BAD_FRAME);
    REQUIRE(produced > 0 && metricBlocks >= 2);
    hwpa_analyzer_destroy(direct);
    hwpa_dev_free(dTraces);
    hwpa_dev_free(dState);
    std::printf("wrapper ok: %u chunks, %d pulses, %d metric blocks forwarded\n", chunks, produced, metricBlocks);
    return 0;
}
