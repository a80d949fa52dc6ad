"""Inputs of the reference's HF-metrics known-answer test
(basecaller/test/TestHFMetricsFilter.cpp:77-245, testPopulated at :249-336):
per batch every ZMW holds ten abutting 4-frame pulses labelled by the
repeating pattern "ACGTGG", and the baseliner / frame-labeler /
pulse-detector metrics and detection models are fixed records."""
import numpy as np

from hw_amd import types as T

FRAMES = 512            # BatchLayoutConfig::framesPerChunk default
BLOCK = 4096            # BasecallerMetricsConfig::framesPerHFMetricBlock default
BATCHES = BLOCK // FRAMES
NUM_BASES = 10
BASE_WIDTH = 4
PATTERN = "ACGTGG"
LABEL = {"A": 0, "C": 1, "G": 2, "T": 3}
MID = {"G": 10, "A": 20, "T": 30, "C": 40}
MAX = {"G": 15, "A": 25, "T": 35, "C": 45}
FRAME_RATE = 100.0


def fixed(v):
    """Pulse::ToFixedPrecision: uint16(v / 0.1f + 0.5f) in single precision."""
    return np.uint16(np.float32(np.float32(v) / np.float32(0.1)) + np.float32(0.5))


def pulses(lanes, batch, max_calls=24, ipd=0):
    """GenerateBases: returns (pulses [lanes*64*max_calls], counts [lanes*64])."""
    p = np.zeros((lanes, 64, max_calls), dtype=T.PULSE)
    for b in range(NUM_BASES):
        ch = PATTERN[b % len(PATTERN)]
        mid = fixed(MID[ch] + b)
        midf = np.float32(mid) * np.float32(0.1)
        mod = np.float32(1.1) if b % 2 == 0 else np.float32(1.01)
        q = p[:, :, b]
        q["label"] = LABEL[ch]
        q["start"] = b * BASE_WIDTH + b * ipd + batch * FRAMES
        q["width"] = BASE_WIDTH
        q["mid_signal"] = mid
        q["mean_signal"] = mid
        q["max_signal"] = fixed(MAX[ch])
        q["signal_m2"] = np.float32(np.float32(midf * midf) * np.float32(BASE_WIDTH - 2)) * mod
        q["is_reject"] = 0
    counts = np.full(lanes * 64, NUM_BASES, dtype=np.uint32)
    return p.reshape(-1), counts


def baseliner_stats(lanes):
    """GenerateBaselineMetrics."""
    s = np.zeros(lanes, dtype=T.BASELINER_STATS)
    for k in range(4):
        s["full"]["fbuf"][:, k, :] = -9.0 + 2 * k
        s["full"]["bbuf"][:, k, :] = 6.0 + k
    s["full"]["bidx"][:, 0, :] = min(FRAMES, 4)
    s["full"]["bidx"][:, 1, :] = FRAMES % 4
    s["raw_baseline_sum"] = 100
    b = s["baseline"]
    b["m0"], b["m1"], b["m2"] = 98, 10, 100
    s["full"]["lag_m2"] = 2510
    f = s["full"]["basic"]
    f["m0"], f["m1"], f["m2"] = FRAMES, 110, 5650
    g = s["background"]
    g["m0"], g["m1"], g["m2"] = 42, 0, 0
    return s


def viterbi_score(lanes):
    """GenerateFrameLabelerMetrics."""
    return np.full(lanes * 64, 1.09, dtype=np.float32)


def pd_stats(lanes):
    """GeneratePulseDetectorMetrics."""
    s = np.zeros(lanes, dtype=T.STAT_ACC)
    s["m0"], s["m1"], s["m2"] = 98, 10, 100
    return s


MODEL_MEANS = (227.13, 154.45, 97.67, 61.32)
MODEL_VARS = (776.0, 426.0, 226.0, 132.0)


def models(lanes):
    """GenerateModels (PBHalf means / variances)."""
    m = np.zeros(lanes, dtype=T.LANE_MODEL)
    h = lambda v: np.float16(v).view(np.uint16)
    for a in range(4):
        m["analogs"]["means"][:, a, :] = h(MODEL_MEANS[a])
        m["analogs"]["vars"][:, a, :] = h(MODEL_VARS[a])
    m["baseline"]["means"] = h(0.0)
    m["baseline"]["vars"] = h(33.0)
    return m


def threshold_tree(feature, threshold, left_value=0, right_value=1):
    """A three-node tree: feature <= threshold -> left_value, else right_value."""
    return dict(children_left=[1, -1, -1], children_right=[2, -1, -1], feature=[feature, -2, -2],
                threshold=[threshold, -2.0, -2.0], value=[0, left_value, right_value])
