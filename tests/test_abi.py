"""The C-ABI library loads, exports every entry point include/hwpa.h
declares, and its structure layouts match the Python mirrors (no GPU)."""
import ctypes
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "hwpa.h")).read()
    return sorted(set(re.findall(r"\b(hwpa_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(H):
    names = _declared()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(H.lib, n)]
    assert not missing, missing


def test_layout_sizes_match(H, T):
    n = H.lib.hwpa_layout_sizes(None, 0)
    arr = (ctypes.c_int64 * n)()
    H.lib.hwpa_layout_sizes(arr, n)
    assert [s for _, s in T.LAYOUT_ORDER] == list(arr)
    # reference record sizes that must hold
    assert T.BASELINER_STATS.itemsize == 6016   # BaselinerStatAccumState
    assert T.LANE_HIST.itemsize == 39168        # LaneHistogram<float, u16>, 300 bins
    assert T.LANE_MODEL.itemsize == 2048        # LaneModelParameters<PBHalf, 64>
    assert T.PULSE.itemsize == 20               # Pulse


def test_config_defaults_and_json(H, T):
    c = H.default_config()
    assert c.filter_type == T.FILTER_TYPES["MultiScaleLarge"]
    assert c.sigma_ema_scale_strides == 512 and c.em_iteration_limit == 20
    assert c.min_frames_for_estimate == 4000 and c.num_frames_pre_accum_stats == 1000
    assert abs(c.bin_size_coeff - 0.2) < 1e-7 and c.max_calls_per_zmw == 48
    H.load_json(c, '{"baselinerConfig": {"Filter": "TwoScaleMedium", "SigmaEmaScaleStrides": 0},'
                   ' "dmeConfig": {"EmIterationLimit": 7, "ConfidenceHalfLife": [100, 200]},'
                   ' "movie": {"analogs": [{"relAmplitude": 3.3, "baseLabel": "T"}]}}')
    assert c.filter_type == 4 and c.sigma_ema_scale_strides == 0 and c.em_iteration_limit == 7
    assert c.confidence_half_life[1] == 200
    assert abs(c.movie.analogs[0].rel_amplitude - 3.3) < 1e-6 and c.movie.analogs[0].base_label == ord("T")


def test_config_rejects_bad_values(H, T):
    c = H.default_config()
    import pytest
    with pytest.raises(H.HwpaError) as e:
        H.load_json(c, '{"pulseAccumConfig": {"XspAmpThresh": 1.5}}')
    assert e.value.code == T.ERR_BAD_CONFIG
    c = H.default_config()
    with pytest.raises(H.HwpaError):
        H.load_json(c, '{"baselinerConfig": {"Filter": "Huge"}}')
    c = H.default_config()
    with pytest.raises(H.HwpaError):
        H.load_json(c, '{"noSuchSection": {"x": 1}}')
    c = H.default_config()
    with pytest.raises(H.HwpaError):
        H.load_json(c, '{"movie": {"analogs": [{"relAmplitude": -1}]}}')


def test_filter_tables(H):
    # traceAnalysis/BaselinerParams.cpp:10-72; LatentSize/AggregateStride BaselinerParams.h:45-70
    exp = {0: ([1, 2, 8], [11, 17, 61], 1024), 1: ([1, 2, 8], [9, 17, 31], 1024),
           2: ([1, 2, 8], [7, 17, 17], 512), 3: ([2, 8], [11, 61], 1024),
           4: ([2, 8], [9, 31], 1024), 5: ([2, 8], [7, 17], 512)}
    # LatentSize = next power of two >= sum_i w_i * prod_{j<=i} s_j, e.g. 9 + 34 + 496 = 539 -> 1024
    for t, (s, w, lat) in exp.items():
        fp = H.filter_params(t)
        assert list(fp.strides)[:fp.num_stages] == s and list(fp.widths)[:fp.num_stages] == w
        assert fp.aggregate_stride == 16 and fp.latent_size == lat


def test_no_gpu_analyzer_fails_loudly(H, T):
    if H.device_count() > 0:
        return
    import pytest
    with pytest.raises(H.HwpaError) as e:
        H.BatchAnalyzer(0, 1, 512, H.bench_config(512))
    assert e.value.code in (6, 5)
