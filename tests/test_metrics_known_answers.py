"""Pins of the oracle's HF-metrics stage and its autocorrelation accumulator
against the reference's own known answers (no GPU):
  common/test/TestAutocorrAccumulator.cpp (one-pass values, merge orders, sine)
  basecaller/test/TestHFMetricsFilter.cpp (testPopulated, :249-336)
and against what the definitions fix (counts on whole sequences, independence
of how the pulse stream is cut into chunks, the activity-tree features read
back through threshold trees)."""
import numpy as np
import pytest

import metrics_inputs as MI


# ---------------------------------------------------------------------------
# AutocorrAccumulator (lag 4)
# ---------------------------------------------------------------------------

def test_autocorr_simple_one_pass(O):
    # SimpleOnePass: samples 0..8 -> autocorrelation -1/3
    assert O.autocorr(np.arange(9)) == pytest.approx(-1.0 / 3, rel=2e-6)


def test_autocorr_square_one_pass(O):
    # SquareOnePass: samples i*i, i < 14 -> 0.22877049
    assert O.autocorr(np.arange(14.0) ** 2) == pytest.approx(0.22877049, rel=2e-6)


def test_autocorr_sine(O):
    # AutocorrSine: sin(0.1 i), i < 100 -> 0.9379457 (atol 5e-5), and the direct sum
    x = np.sin(np.float32(0.1) * np.arange(100, dtype=np.float32)).astype(np.float32)
    r = O.autocorr(x)
    assert abs(r - 0.9379457) < 5e-5
    xd = x.astype(np.float64)
    d = xd - xd.mean()
    direct = (d[:-4] * d[4:]).sum() / (100 - 4 - 1) / xd.var(ddof=1)
    assert abs(r - direct) < 5e-5


def _acc_close(a, b):
    # AccCompare: |a - b| <= atol + rtol |b|, rtol 1e-5, atol 1e-8
    return abs(a - b) <= 1e-8 + 1e-5 * abs(b)


def test_autocorr_simple_and_partial_merge(O):
    # SimpleMerge: 0..29 in three parts of 10; PartialMerge: 1..30 in parts 1, 7, 19, 3
    x = np.arange(30, dtype=np.float32)
    assert _acc_close(O.autocorr(x), O.autocorr(x, [10, 10, 10]))
    assert _acc_close(O.autocorr(x + 1), O.autocorr(x + 1, [1, 7, 19, 3]))


@pytest.mark.parametrize("nom,den", [(n, d) for n in (2, 4, 7) for d in (8, 12, 20)])
def test_autocorr_merge_threshold(O, nom, den):
    # TestAutocorrAccumulatorLinspace.MergeThreshold: sin(0.6 i), i < 1080, cut at 1080 nom / den
    x = np.sin(np.float32(0.6) * np.arange(1080, dtype=np.float32)).astype(np.float32)
    cut = 1080 * nom // den
    assert _acc_close(O.autocorr(x), O.autocorr(x, [cut, 1080 - cut]))
    assert _acc_close(O.autocorr(x), O.autocorr(x, 6))


def test_autocorr_too_few_samples_is_nan(O):
    # SimpleOnePass: a single sample has no autocorrelation
    assert np.isnan(O.autocorr(np.zeros(1)))


# ---------------------------------------------------------------------------
# HFMetricsFilter testPopulated
# ---------------------------------------------------------------------------

def _run_block(O, H, T, lanes=2, tree=None, chunks=None):
    cfg = H.default_config()
    assert cfg.frames_per_hf_metric_block == MI.BLOCK and cfg.realtime_activity_labels == 1
    cfg.movie.frame_rate = MI.FRAME_RATE
    m = O.Metrics(cfg, lanes)
    if tree is not None:
        m.set_activity_tree(tree[0])
    done = []
    out = None
    for b in range(MI.BATCHES):
        p, c = MI.pulses(lanes, b)
        d, o = m.process(p, c, 24, MI.FRAMES, (b + 1) * MI.FRAMES, MI.baseliner_stats(lanes),
                         MI.viterbi_score(lanes), MI.pd_stats(lanes), MI.models(lanes), T.METRICS)
        done.append(d)
        if d:
            out = o
    assert done == [0] * (MI.BATCHES - 1) + [1]
    return out


def test_hf_metrics_populated(O, H, T):
    mb = _run_block(O, H, T)
    nb = MI.BATCHES
    eq = lambda f, v: np.testing.assert_array_equal(mb[f], np.broadcast_to(v, mb[f].shape), err_msg=f)
    near = lambda f, v, tol: np.testing.assert_allclose(mb[f], np.broadcast_to(np.float32(v), mb[f].shape),
                                                        rtol=0, atol=tol, err_msg=f)
    eq("num_pulse_frames", nb * MI.NUM_BASES * MI.BASE_WIDTH)
    eq("num_base_frames", nb * MI.NUM_BASES * MI.BASE_WIDTH)
    eq("num_half_sandwiches", nb * (MI.NUM_BASES - 2))
    eq("num_sandwiches", nb)
    eq("num_pulse_label_stutters", nb)
    per_analog = np.array([2, 2, 4, 2])[None, :, None] * nb
    eq("num_pulses_by_analog", per_analog)
    eq("num_bases_by_analog", per_analog)
    eq("num_pulses", nb * MI.NUM_BASES)
    eq("num_bases", nb * MI.NUM_BASES)
    eq("start_frame", 0)
    eq("num_frames", nb * MI.FRAMES)
    eq("pk_mid_signal", np.array([736, 1408, 944, 1152], np.float32)[None, :, None])
    eq("pk_max", np.array([25, 45, 15, 35], np.float32)[None, :, None])
    for f, vals in (("bp_zvar", (0.0171916, 0.0043878, 0.0192236, 0.0065546)),
                    ("pk_zvar", (0.0706076, 0.0451073, 0.0744171, 0.0960086))):
        for a in range(4):
            np.testing.assert_allclose(mb[f][:, a, :], vals[a], rtol=0, atol=1e-3, err_msg=f)
    eq("num_pk_mid_frames", per_analog * (MI.BASE_WIDTH - 2))
    eq("num_pk_mid_bases_by_analog", per_analog)
    near("autocorrelation", 0.415835, 1e-3)
    near("pulse_detection_score", 0.002128, 1e-4)
    near("num_frames_baseline", 98 * nb, 1e-4)
    near("frame_baseline_dws", 10 / 98, 1e-4)
    near("frame_baseline_variance_dws", (100 - 10 * 10 / 98) / (98 - 1), 1e-2)
    eq("pixel_checksum", 0)
    eq("activity_label", 0)     # no tree supplied: labels stay EMPTY


def _ordered(f):
    i = np.float32(f).view(np.int32).astype(np.int64)
    return i if i >= 0 else -(i & 0x7FFFFFFF)


def _unordered(i):
    i = int(i)
    return np.int32(i).view(np.float32) if i >= 0 else np.int32((-i) | -0x80000000).view(np.float32)


def _feature_through_tree(O, H, T, k, lo=-1e4, hi=1e4):
    """The smallest float32 threshold t with feature_k <= t, found by bisecting
    over threshold trees; equals feature_k itself (ZMW 0 of lane 0)."""
    a, b = _ordered(lo), _ordered(hi)
    while b - a > 1:
        mid = (a + b) // 2
        tree = H.activity_tree(**MI.threshold_tree(k, _unordered(mid)), hsw_curve=(0.0, 0.0),
                               max_acceptable_hsw_rate=1.0)
        lab = _run_block(O, H, T, lanes=1, tree=tree)["activity_label"][0, 0]
        if lab == 0:
            b = mid
        else:
            a = mid
    return float(_unordered(b))


def test_activity_features_through_trees(O, H, T):
    # The eleven ActivityLabeler features of the populated block, from their
    # definitions: rates per pulse, seconds = frames / frame rate, amplitude
    # ratios of the model means, the lowest-amplitude analog left out of the
    # zvar averages.
    mb = _run_block(O, H, T, lanes=1)
    n_p = MI.BATCHES * MI.NUM_BASES
    seconds = MI.BATCHES * MI.FRAMES / MI.FRAME_RATE
    blvar = float(mb["frame_baseline_variance_dws"][0, 0])
    sd = np.sqrt(blvar)
    pkmid = mb["pk_mid_signal"][0, :, 0] / mb["num_pk_mid_frames"][0, :, 0]
    means = np.float16(MI.MODEL_MEANS).astype(np.float64)
    nba = mb["num_bases_by_analog"][0, :, 0]
    expect = {
        0: n_p / seconds,
        1: MI.BATCHES / n_p,
        2: MI.BATCHES * (MI.NUM_BASES - 2) / n_p,       # hsw curve 0 -> expected rate 0
        3: 1.09 / MI.FRAMES,
        4: MI.BASE_WIDTH,
        5: MI.BATCHES / n_p,
        6: sum(pkmid[a] / sd * nba[a] / n_p * means.min() / means[a] for a in range(4)),
        7: max((mb["pk_max"][0, a, 0] - pkmid[a]) / sd for a in range(4)),
        8: float(mb["autocorrelation"][0, 0]),
        9: float(np.mean(mb["bp_zvar"][0, :3, 0])),     # analog 3 (61.32) has the lowest amplitude
        10: float(np.mean(mb["pk_zvar"][0, :3, 0])),
    }
    for k, v in expect.items():
        got = _feature_through_tree(O, H, T, k)
        assert got == pytest.approx(v, rel=2e-5, abs=1e-7), k


# ---------------------------------------------------------------------------
# Counting over random pulse streams, cut into chunks at random places
# ---------------------------------------------------------------------------

def _whole_sequence_counts(seq):
    """Stutters, half sandwiches and sandwiches of one ZMW's whole pulse list
    [(start, width, label)], straight from their definitions."""
    st = hs = sw = 0
    for i in range(1, len(seq)):
        (s0, w0, l0), (s1, w1, l1) = seq[i - 1], seq[i]
        abut = s1 == s0 + w0
        st += l1 == l0
        hs += abut and l1 != l0
        if i >= 2:
            sm, wm, lm = seq[i - 2]
            sw += abut and s0 == sm + wm and lm == l1 and lm != l0
    return st, hs, sw


@pytest.mark.parametrize("seed", [1, 2])
def test_counts_independent_of_chunking(O, H, T, seed):
    rng = np.random.default_rng(seed)
    lanes, max_calls, frames = 1, 40, 64
    nch = 6
    seqs = []
    for z in range(64):
        t, seq = 0, []
        while True:
            t += int(rng.integers(0, 3))                 # ipd 0 abuts
            w = int(rng.integers(1, 6))
            if t + w > nch * frames:
                break
            seq.append((t, w, int(rng.integers(0, 4))))
            t += w
        seqs.append(seq)
    cfg = H.default_config()
    cfg.frames_per_hf_metric_block = nch * frames
    m = O.Metrics(cfg, lanes)
    st = np.zeros(1, T.BASELINER_STATS)
    st["full"]["basic"]["m0"] = frames
    out = None
    for c in range(nch):
        p = np.zeros((64, max_calls), dtype=T.PULSE)
        counts = np.zeros(64, np.uint32)
        for z, seq in enumerate(seqs):
            mine = [q for q in seq if c * frames <= q[0] < (c + 1) * frames]
            assert len(mine) <= max_calls
            for j, (s, w, lab) in enumerate(mine):
                p[z, j]["start"], p[z, j]["width"], p[z, j]["label"] = s, w, lab
                p[z, j]["mid_signal"] = p[z, j]["max_signal"] = 100
            counts[z] = len(mine)
        d, o = m.process(p.reshape(-1), counts, max_calls, frames, (c + 1) * frames, st,
                         np.zeros(64, np.float32), np.zeros(1, T.STAT_ACC), MI.models(1), T.METRICS)
        if d:
            out = o
    assert out is not None
    for z, seq in enumerate(seqs):
        s, h, w = _whole_sequence_counts(seq)
        assert out["num_pulse_label_stutters"][0, z] == s
        assert out["num_half_sandwiches"][0, z] == h
        assert out["num_sandwiches"][0, z] == w
        assert out["num_pulses"][0, z] == len(seq)
        assert out["num_pulse_frames"][0, z] == sum(q[1] for q in seq)
