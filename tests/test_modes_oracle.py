"""The analyzer's model-estimation modes and trace encodings in the oracle
(no GPU): Fixed (FixedModelBatchAnalyzer with StaticDetModelConfig),
Initial (SingleEstimateBatchAnalyzer), Dynamic (default), and uint8 traces,
which the reference zero-extends to int16 before any arithmetic
(dataTypes/BatchData.h:160-182)."""
import numpy as np
import pytest

from conftest import gen_traces


def _dts(T):
    return (T.PULSE, T.STAT_ACC, T.BASELINER_STATS, T.LANE_MODEL)


def test_mode_config_json(H, T):
    c = H.default_config()
    assert c.model_estimation_mode == T.MODEL_DYNAMIC
    assert c.static_baseline_mean == 0.0 and c.static_baseline_variance == 33.0
    H.load_json(c, '{"modelEstimationMode": "FixedEstimations",'
                   ' "staticDetModelConfig": {"baselineMean": 2.5, "baselineVariance": 40}}')
    assert c.model_estimation_mode == T.MODEL_FIXED
    assert c.static_baseline_mean == 2.5 and c.static_baseline_variance == 40.0
    H.load_json(c, '{"modelEstimationMode": "InitialEstimations"}')
    assert c.model_estimation_mode == T.MODEL_INITIAL
    for bad in ('{"modelEstimationMode": "Sometimes"}', '{"staticDetModelConfig": {"baselineVariance": -1}}'):
        with pytest.raises(H.HwpaError) as e:
            H.load_json(H.default_config(), bad)
        assert e.value.code == T.ERR_BAD_CONFIG


def test_fixed_model_closed_form(H, O, T):
    # StaticDetModelConfig::SetupAnalogs with the default movie (refSnr 50/12, relative
    # amplitudes 3.6/2.4/1.6/1.0, excess noise CV 0.1) and baseline N(0, 33):
    # refSignal = 50/12 * sqrt(33) = 23.9357; mean = rel * refSignal; var = 33 + mean + (0.1 mean)^2
    cfg = H.bench_config(512)
    cfg.model_estimation_mode = T.MODEL_FIXED
    L, F = 2, 512
    oa = O.Analyzer(cfg, 0, L, F)
    r = oa(gen_traces(O, T, H, L, F, 1, seed=5)[0], 0, _dts(T), want=("models", "pulses"))
    m = r["models"]
    ref = 50.0 / 12.0 * np.sqrt(33.0)
    for a, rel in enumerate((3.6, 2.4, 1.6, 1.0)):
        mean = rel * ref
        var = 33.0 + mean + (0.1 * mean) ** 2
        np.testing.assert_allclose(T.half_to_float(m["analogs"]["means"][:, a, :]), mean, rtol=1e-3)
        np.testing.assert_allclose(T.half_to_float(m["analogs"]["vars"][:, a, :]), var, rtol=1e-3)
    assert np.all(T.half_to_float(m["baseline"]["means"]) == 0.0)
    assert np.all(T.half_to_float(m["baseline"]["vars"]) == 33.0)
    assert r["flags"][0] == 1 and r["flags"][2] == 1 and r["flags"][1] == 0   # pulses from the first chunk
    assert r["counts"].sum() > 0


def test_fixed_model_runs_every_stage_every_chunk(H, O, T):
    cfg = H.bench_config(512)
    cfg.model_estimation_mode = T.MODEL_FIXED
    cfg.frames_per_hf_metric_block = 1024
    L, F = 2, 512
    oa = O.Analyzer(cfg, 3, L, F)                     # no pool stagger in this mode
    flags = [oa(tr, c * F, _dts(T), want=("models",))["flags"]
             for c, tr in enumerate(gen_traces(O, T, H, L, F, 5, seed=6))]
    assert [f[0] for f in flags] == [1] * 5 and [f[2] for f in flags] == [1] * 5
    assert [f[1] for f in flags] == [0] * 5
    # pulse batches end at 512 c + 488: a 1024-frame block ends in chunks 0, 2, 4
    assert [f[3] for f in flags] == [1, 0, 1, 0, 1]


def test_initial_mode_keeps_the_first_estimate(H, O, T):
    # Measurement configuration: baseliner start-up 1024 frames, one estimate per 512.
    # SingleEstimate feeds the DME while first_frame > 1024: chunk 3 (histogram init),
    # chunk 4 (first estimate); the model never changes afterwards.
    cfg = H.bench_config(512)
    cfg.model_estimation_mode = T.MODEL_INITIAL
    cfg.frames_per_hf_metric_block = 1024
    L, F, C = 2, 512, 8
    oa = O.Analyzer(cfg, 1, L, F)
    res = [oa(tr, c * F, _dts(T), want=("models", "pulses")) for c, tr in enumerate(gen_traces(O, T, H, L, F, C, 9))]
    assert [r["flags"][1] for r in res] == [0, 0, 0, 0, 1, 0, 0, 0]
    assert [r["flags"][0] for r in res] == [0, 0, 0, 0, 1, 1, 1, 1]
    assert [r["flags"][2] for r in res] == [0, 0, 0, 0, 1, 1, 1, 1]
    assert [r["flags"][3] for r in res] == [1, 0, 1, 0, 1, 0, 1, 0]   # metrics run from the first chunk
    for r in res[:4]:
        assert r["counts"].sum() == 0
    for r in res[5:]:
        assert r["models"].tobytes() == res[4]["models"].tobytes()
    assert sum(int(r["counts"].sum()) for r in res[4:]) > 0


def test_dynamic_mode_reestimates(H, O, T):
    cfg = H.bench_config(512)
    L, F, C = 2, 512, 7
    oa = O.Analyzer(cfg, 0, L, F)
    res = [oa(tr, c * F, _dts(T), want=("models",)) for c, tr in enumerate(gen_traces(O, T, H, L, F, C, 9))]
    est = [r["flags"][1] for r in res]
    assert sum(est) >= 2
    i, j = [k for k, e in enumerate(est) if e][:2]
    assert res[i]["models"].tobytes() != res[j]["models"].tobytes()


def _u8_chunks(O, T, H, L, F, C, seed):
    # 8-bit traces: halved, offset by a pedestal of 64, clipped to the byte range
    return [np.clip(tr.astype(np.int32) // 2 + 64, 0, 255).astype(np.uint8) for tr in gen_traces(O, T, H, L, F, C, seed)]


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_uint8_traces_equal_zero_extended_int16(H, O, T, mode):
    cfg = H.bench_config(512)
    cfg.model_estimation_mode = mode
    cfg.movie.pedestal = 64
    L, F, C = 2, 512, 6
    a8, a16 = O.Analyzer(cfg, 0, L, F), O.Analyzer(cfg, 0, L, F)
    want = ("pulses", "labels", "models", "baselined", "metrics")
    for c, u in enumerate(_u8_chunks(O, T, H, L, F, C, 12)):
        r8 = a8(u, c * F, _dts(T), want=want)
        r16 = a16(u.astype(np.int16), c * F, _dts(T), want=want)
        assert r8["flags"] == r16["flags"]
        for k in ("baselined", "labels", "counts"):
            assert np.array_equal(r8[k], r16[k]), (c, k)
        for k in ("pulses", "models", "metrics"):
            assert r8[k].tobytes() == r16[k].tobytes(), (c, k)
