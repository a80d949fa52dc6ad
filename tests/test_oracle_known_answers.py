"""Pins of the CPU oracle against what the reference's own tests and the
mathematics fix (no GPU).  Each test names the reference test or the closed
form it restates; a plausible slip (wrong index, sign, window edge, order)
in the restatement fails one of them."""
import ctypes
import itertools

import numpy as np
import pytest

from conftest import gen_traces

_vp = ctypes.c_void_p


def _p(a):
    return a.ctypes.data_as(_vp)


# ---------------------------------------------------------------------------
# Extremum (erode/dilate) filter -- PBF/test/testExtremaFilter.cu:21-41 and
# the closed form out(t) = op(x[t-W .. t-1]) with W zeros of history.
# ---------------------------------------------------------------------------

def _hgw(O, width, stride, is_max, x):
    chunks, length = x.shape[0], x.shape[1]
    nout = (length + stride - 1) // stride
    out = np.empty((chunks, nout, 64), dtype=np.int16)
    O.lib.orc_hgw(ctypes.c_int32(width), ctypes.c_int32(stride), ctypes.c_int32(1 if is_max else 0),
                  ctypes.c_int32(chunks), ctypes.c_int32(length), _p(np.ascontiguousarray(x)), _p(out))
    return out


def test_hgw_sawtooth_known_answer(O):
    # sawtooth of period 25 through a width-7 max filter: out(t) = (t-1) mod 25,
    # or 24 while the window still holds the previous period's top value.
    H, W = 25, 7
    t = np.arange(400)
    x = np.broadcast_to((t % H).astype(np.int16)[None, :, None], (1, 400, 64)).copy()
    out = _hgw(O, W, 1, True, x)[0, :, 0].astype(int)
    for tt in range(W, 400):
        e = (tt - 1) % H
        if e < W - 1:
            e = H - 1
        assert out[tt] == e, tt


@pytest.mark.parametrize("width,stride,is_max", [(11, 1, False), (11, 1, True), (17, 2, False), (61, 8, True),
                                                 (7, 2, True), (31, 8, False)])
def test_hgw_matches_sliding_window(O, width, stride, is_max):
    rng = np.random.default_rng(width * 10 + stride)
    chunks, length = 5, 256
    x = rng.integers(-3000, 3000, size=(chunks, length, 64)).astype(np.int16)
    out = _hgw(O, width, stride, is_max, x)
    dec = x[:, ::stride, :].reshape(-1, 64)                 # decimated stream, chunk phase 0
    hist = np.concatenate([np.zeros((width, 64), np.int16), dec])
    op = np.max if is_max else np.min
    ref = np.stack([op(hist[k:k + width], axis=0) for k in range(dec.shape[0])])
    assert np.array_equal(out.reshape(-1, 64), ref)


# ---------------------------------------------------------------------------
# Baseliner -- closed forms on constant input (BT/TestBaseliner.cpp:483-577)
# and the latent-frame bookkeeping of the statistics.
# ---------------------------------------------------------------------------

def test_baseliner_constant_input_closed_form(H, O, T):
    cfg = H.default_config(sigma_ema_scale_strides=0.0)     # sigma EMA off -> sigma = sqrt(1/12)
    L, F = 2, 512
    c = np.int16(137)
    bl = O.Baseliner(cfg, L)
    outs = [bl.run(np.full((L, F, 64), c, np.int16), T.BASELINER_STATS) for _ in range(6)]
    # once the zero history has left both windows of the opening (erode then
    # dilate, 2 x (11 + 2*17 + 16*61) frames) the envelopes equal c and
    # blEst = c + 0.78 * 0.2887 -> round(c - blEst) = 0
    assert outs[0][0][0, 0, 0] == c          # frame 0: envelopes still hold the zero history
    for out, st in outs[4:]:
        assert (out == 0).all()
        assert np.all(st["baseline"]["m0"] == F)             # every frame is baseline
        assert np.all(st["baseline"]["m1"] == 0) and np.all(st["baseline"]["m2"] == 0)
        assert np.all(st["trace_min"] == 0) and np.all(st["trace_max"] == 0)
        # background = rounded subtracted baseline, once per 16-frame stride block
        assert np.all(st["background"]["m0"] == F // 16)
        assert np.all(st["background"]["m1"] == (F // 16) * int(c))
        # raw baseline sum accumulates raw (scaled, rounded) frames flagged baseline
        assert np.all(st["raw_baseline_sum"] == F * int(c))


def test_baseliner_stats_follow_latent_frame(H, O, T):
    # Statistics are accumulated one frame late: chunk k's moments cover its
    # frames 0..F-2 plus the previous chunk's last frame (zero initially).
    cfg = H.default_config()
    L, F = 2, 256
    chunks = gen_traces(O, T, H, L, F, 3, seed=4)
    bl = O.Baseliner(cfg, L)
    prev_last = np.zeros((L, 64), np.float64)
    for tr in chunks:
        out, st = bl.run(tr, T.BASELINER_STATS)
        o = out.astype(np.float64)
        seq = np.concatenate([prev_last[:, None, :], o[:, :-1, :]], axis=1)   # [L, F, 64]
        assert np.all(st["full"]["basic"]["m0"] == F)
        np.testing.assert_allclose(st["full"]["basic"]["m1"], seq.sum(axis=1), rtol=1e-5, atol=1e-2)
        np.testing.assert_allclose(st["full"]["basic"]["m2"], (seq ** 2).sum(axis=1), rtol=1e-5)
        np.testing.assert_allclose(st["full"]["lag_m2"], (seq[:, 4:] * seq[:, :-4]).sum(axis=1), rtol=1e-5, atol=1)
        assert np.array_equal(st["trace_min"], seq.min(axis=1).astype(np.int16))
        assert np.array_equal(st["trace_max"], seq.max(axis=1).astype(np.int16))
        assert np.array_equal(st["full"]["fbuf"], np.transpose(seq[:, :4, :], (0, 1, 2)).astype(np.float32))
        prev_last = o[:, -1, :]


def test_stats_merge_equals_concatenated_stream(H, O, T):
    # AutocorrAccumulator::Merge (common/AutocorrAccumulator.cpp:72-104): the
    # merged lag-4 moment equals that of the concatenated frame sequence.
    cfg = H.default_config()
    L, F = 2, 256
    bl = O.Baseliner(cfg, L)
    agg = O.Aggregator(L)
    seqs = []
    prev_last = np.zeros((L, 64))
    for tr in gen_traces(O, T, H, L, F, 4, seed=8):
        out, st = bl.run(tr, T.BASELINER_STATS)
        o = out.astype(np.float64)
        seqs.append(np.concatenate([prev_last[:, None, :], o[:, :-1, :]], axis=1))
        prev_last = o[:, -1, :]
        agg.add(st)
    s = agg.get(T.BASELINER_STATS)
    full = np.concatenate(seqs, axis=1)
    assert np.all(s["full"]["basic"]["m0"] == full.shape[1])
    np.testing.assert_allclose(s["full"]["lag_m2"], (full[:, 4:] * full[:, :-4]).sum(axis=1), rtol=2e-5, atol=1)
    np.testing.assert_allclose(s["full"]["basic"]["m2"], (full ** 2).sum(axis=1), rtol=2e-5)
    assert np.array_equal(s["full"]["fbuf"], full[:, :4, :].astype(np.float32))
    assert np.array_equal(s["full"]["bbuf"][:, [0, 1, 2, 3], :], full[:, -4:, :].astype(np.float32))
    assert np.array_equal(s["trace_min"], full.min(axis=1).astype(np.int16))


# ---------------------------------------------------------------------------
# Histogram -- BT/TestTraceHistogramAccumulator.cpp:325-465
# ---------------------------------------------------------------------------

def _stats_with_baseline(T, lanes, m0, m1, m2):
    s = np.zeros(lanes, dtype=T.BASELINER_STATS)
    s["baseline"]["m0"] = m0
    s["baseline"]["m1"] = m1
    s["baseline"]["m2"] = m2
    s["full"]["basic"]["m0"] = np.maximum(m0, 1)
    return s


def test_histogram_reset_from_stats(H, O, T):
    # ResetFromStats (:325-381): m0 = t, m1 = t^2, m2 = 2 t^3 -> mean t,
    # var t^3/(t-1); below BaselineStatMinFrameCount the fallback sigma is used.
    cfg = H.default_config()
    L = 2
    tmp = (np.arange(64)[None, :] + 64 * np.arange(L)[:, None] + 2).astype(np.float32)
    s = _stats_with_baseline(T, L, tmp, tmp * tmp, 2 * tmp * tmp * tmp)
    hist = O.Histogram(cfg, L)
    assert hist.reset(s) == 0
    h = hist.get(T.LANE_HIST)
    sig = np.where(tmp >= cfg.baseline_stat_min_frame_count, np.sqrt(tmp.astype(np.float64) ** 3 / (tmp - 1.0)),
                   cfg.fallback_baseline_sigma)
    exp_bin = np.maximum(cfg.bin_size_coeff * sig, 1.0)
    np.testing.assert_allclose(h["bin_size"], exp_bin, rtol=4 * 1.2e-7)
    assert np.all(h["low_bound"] < tmp - 2 * sig)


def _unit_bins_hist(O, T, cfg, lanes):
    # stats giving mean 20, sigma 5 -> binSize 1, lowBound 0: bin i holds value i
    s = _stats_with_baseline(T, lanes, np.full((lanes, 64), 100, np.float32), np.full((lanes, 64), 2000, np.float32),
                             np.full((lanes, 64), 42475, np.float32))
    hist = O.Histogram(cfg, lanes)
    assert hist.reset(s) == 0
    h = hist.get(T.LANE_HIST)
    assert np.all(h["low_bound"] == 0) and np.all(h["bin_size"] == 1)
    return hist, O.dme_init_models(cfg, s, T.LANE_MODEL)


def test_histogram_constant_lanes(H, O, T):
    # SingleLaneConstant / MultiLaneConstant (:383-465): ZMW j of lane l holds
    # 64 l + j; all frames but the first two (edge classifier warm-up) land in
    # that bin; no outliers.
    cfg = H.default_config()
    L, F, nblocks = 4, 512, 4
    hist, models = _unit_bins_hist(O, T, cfg, L)
    val = (np.arange(64)[None, :] + 64 * np.arange(L)[:, None]).astype(np.int16)
    for _ in range(nblocks):
        hist.add(np.broadcast_to(val[:, None, :], (L, F, 64)).copy(), models)
    h = hist.get(T.LANE_HIST)
    assert np.all(h["outlier_low"] == 0) and np.all(h["outlier_high"] == 0)
    for l in range(L):
        for j in range(64):
            counts = h["bin_count"][l, :, j]
            assert counts[64 * l + j] == F * nblocks - 2
            assert counts.sum() == F * nblocks - 2


def test_histogram_outliers(H, O, T):
    cfg = H.default_config()
    L, F = 1, 256
    hist, models = _unit_bins_hist(O, T, cfg, L)
    tr = np.zeros((L, F, 64), np.int16)
    tr[:, :, :16] = -5       # below lowBound
    tr[:, :, 16:32] = 300    # at upper bound -> high outlier
    tr[:, :, 32:] = 299
    hist.add(tr, models)
    h = hist.get(T.LANE_HIST)
    assert np.all(h["outlier_low"][0, :16] == F - 2) and np.all(h["outlier_high"][0, 16:32] == F - 2)
    assert np.all(h["bin_count"][0, 299, 32:] == F - 2)


# ---------------------------------------------------------------------------
# Pulse accumulator -- BT/TestPulseAccumulator.cpp:194-355, restated with the
# 24-frame latent window of this path (ipd 8 / pw 8 keep pulses out of the
# latent/current seam).
# ---------------------------------------------------------------------------

def test_pulse_accumulator_known_pattern(H, O, T):
    cfg = H.default_config()
    L, F, NL = 1, 512, 24
    ipd, pw = 8, 8
    first = 512
    rng = np.random.default_rng(0)
    labels, trace = [], []
    base = 0
    while len(labels) < F:
        labels += [0] * ipd
        trace += list(np.round(rng.normal(0, 10, ipd)))
        a = base % 4
        labels += [a + 5] + [a + 1] * (pw - 2) + [a + 9]
        fstart = len(labels) - pw
        trace += [40 if fstart < NL else 50] * pw
        base += 1
    lab = np.broadcast_to(np.int16(labels)[None, :, None], (L, F, 64)).copy()
    trc = np.broadcast_to(np.int16(trace)[None, :, None], (L, F, 64)).copy()
    latent = trc[:, :NL, :].copy()
    cur = np.zeros((L, F, 64), np.int16)
    cur[:, :F - NL] = trc[:, NL:]
    models = np.zeros(L, dtype=T.LANE_MODEL)
    pulser = O.Pulser(cfg, L)
    p, counts, ovf, bls = pulser.run(lab, latent, cur, first - NL, models, T.PULSE, T.STAT_ACC)
    nper = F // (ipd + pw)
    assert np.all(counts == nper - 1) and np.all(ovf == 0)   # the last pulse closes in the next chunk
    for z in (0, 17, 63):
        for k in range(nper - 1):
            q = p[z, k]
            start = (first - NL) + k * (ipd + pw) + ipd
            v = 40 if start < first else 50
            assert q["start"] == start and q["width"] == pw
            assert q["label"] == k % 4                       # A, C, G, T
            assert q["mean_signal"] == v * 10 and q["mid_signal"] == v * 10 and q["max_signal"] == v * 10
            assert abs(q["signal_m2"] - v * v * (pw - 2)) <= v
            assert q["is_reject"] == 0
    baseline_frames = 1 + (ipd - 1) * nper
    assert np.all(bls["m0"] == baseline_frames)
    mean = bls["m1"] / bls["m0"]
    assert np.all(np.abs(mean) < 20)


def test_pulse_fixed_point():
    # Pulse::ToFixedPrecision (dataTypes/Pulse.h:212-222): u16(v / 0.1f + 0.5f)
    v = np.float32([0.0, 0.04, 0.05, 12.34, 6553.5])
    fixed = (v / np.float32(0.1) + np.float32(0.5)).astype(np.uint16)
    assert list(fixed) == [0, 0, 1, 123, 65535]


# ---------------------------------------------------------------------------
# Transition matrix -- each state's outgoing probabilities sum to one
# (construction of traceAnalysis/SubframeScorer.cpp:147-211).
# ---------------------------------------------------------------------------

def test_transition_columns_sum_to_one(H, O, T):
    for kw in ({}, {"viterbi_alpha": 0.8, "viterbi_beta": 1.1, "viterbi_gamma": 0.9}):
        cfg = H.default_config(**kw)
        tr = O.transition_matrix(cfg, T.Transition())
        lp = np.array(tr.logp, dtype=np.float64)
        mask = np.zeros((13, 13), bool)
        for r in range(13):
            mask[r, tr.first_col[r]] = True
            mask[r, tr.seg_first[r]:tr.seg_last[r]] = True
        assert mask.sum() == 61
        p = np.where(mask, np.exp(lp), 0.0)
        assert np.all(lp[mask] < 0)
        np.testing.assert_allclose(p.sum(axis=0), np.ones(13), atol=2e-6)
        # the library computes the same matrix
        tr2 = H.transition_matrix(cfg)
        assert np.array_equal(np.array(tr2.logp), np.array(tr.logp))


# ---------------------------------------------------------------------------
# Viterbi recursion -- exhaustive path enumeration on tiny inputs.
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("seed", [0, 1, 2])
def test_viterbi_recursion_brute_force(H, O, T, seed):
    cfg = H.default_config()
    tr = O.transition_matrix(cfg, T.Transition())
    lp = np.array(tr.logp, dtype=np.float64)
    allowed = np.zeros((13, 13), bool)
    for r in range(13):
        allowed[r, tr.first_col[r]] = True
        allowed[r, tr.seg_first[r]:tr.seg_last[r]] = True
    rng = np.random.default_rng(seed)
    n = 5
    scores = rng.normal(-3, 2, size=(n, 13)).astype(np.float32)
    roi = (rng.random(n) < 0.7).astype(np.uint8)
    start = int(rng.integers(0, 13))
    ll0 = np.full(13, -np.inf, np.float32)
    ll0[start] = 0.0
    path = np.empty(n, np.int16)
    llout = np.empty(13, np.float32)
    O.lib.orc_viterbi_core(ctypes.byref(cfg), ctypes.c_int32(n), _p(scores), _p(roi), _p(ll0), _p(path), _p(llout))
    best, best_path = -np.inf, None
    for states in itertools.product(range(13), repeat=n):
        prev, tot, ok = start, 0.0, True
        for f, s in enumerate(states):
            if not allowed[s, prev] or (s != 0 and not roi[f]):
                ok = False
                break
            tot += scores[f, s] + lp[s, prev]
            prev = s
        if ok and tot > best:
            best, best_path = tot, states
    assert tuple(int(x) for x in path) == best_path
    assert abs(float(llout.max()) - best) < 1e-3


# ---------------------------------------------------------------------------
# EM -- statistical recovery (BT/TestDmeEm.cpp:160-244) and invariants.
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("form", ["reference", "mirror"])
@pytest.mark.parametrize("snr", [4.0, 6.0])
def test_em_recovers_mixture(H, O, T, snr, form):
    cfg = H.default_config()
    L = 1
    rng = np.random.default_rng(int(snr))
    sigma = 10.0
    amp1 = snr * sigma                    # dimmest analog signal
    rel = np.array([3.6, 2.4, 1.6, 1.0])
    means = np.concatenate([[0.0], amp1 * rel])
    cv2 = 0.1 ** 2
    var = np.concatenate([[sigma ** 2], sigma ** 2 + means[1:] + cv2 * means[1:] ** 2])
    w = np.array([0.6, 0.1, 0.1, 0.1, 0.1])
    nframes = 4000
    hist = np.zeros(L, dtype=T.LANE_HIST)
    stats = np.zeros(L, dtype=T.BASELINER_STATS)
    lo, bs = -4 * sigma, 0.2 * sigma
    for z in range(64):
        comp = rng.choice(5, size=nframes, p=w)
        x = rng.normal(means[comp], np.sqrt(var[comp]))
        b = np.floor((x - lo) / bs).astype(int)
        hist["low_bound"][0, z], hist["bin_size"][0, z] = lo, bs
        hist["outlier_low"][0, z] = (b < 0).sum()
        hist["outlier_high"][0, z] = (b >= 300).sum()
        np.add.at(hist["bin_count"][0, :, z], b[(b >= 0) & (b < 300)], 1)
        xb = x[comp == 0]
        stats["baseline"]["m0"][0, z] = len(xb)
        stats["baseline"]["m1"][0, z] = xb.sum()
        stats["baseline"]["m2"][0, z] = (xb ** 2).sum()
        stats["full"]["basic"]["m0"][0, z] = nframes
    cfg.movie.ref_snr = snr
    models = O.dme_init_models(cfg, stats, T.LANE_MODEL)
    out, diag = O.dme_estimate(cfg, hist, stats, models, (0, 4000), (0, 4000), want_diag=True, form=form)
    est = np.array(diag[0].est_fp32)          # [mode][w, mean, var][zmw]
    assert np.mean(np.array(diag[0].converged)) > 0.9
    np.testing.assert_allclose(est[1:, 1, :].mean(axis=1), means[1:], rtol=0.05)
    assert abs(est[0, 1, :].mean()) < 0.1 * sigma
    np.testing.assert_allclose(np.sqrt(est[0, 2, :].mean()), sigma, rtol=0.1)
    np.testing.assert_allclose(est[:, 0, :].mean(axis=1), w, atol=0.03)
    # invariants: weights sum to 1, analog means are one scale times the relative amplitudes
    np.testing.assert_allclose(est[:, 0, :].sum(axis=0), 1.0, rtol=1e-5)
    ratio = est[1:, 1, :] / rel[:, None]
    np.testing.assert_allclose(ratio, np.broadcast_to(ratio[0], ratio.shape), rtol=1e-6)
    assert np.all(est[0, 2, :] >= 1.0 / 12.0)


@pytest.mark.parametrize("form", ["reference", "mirror"])
def test_em_invariants_on_pipeline_data(H, O, T, form):
    from test_gpu_parity import _em_case
    L, F = 2, 512
    cfg, h, s, m = _em_case(H, O, T, L, F, seed=5)
    out, diag = O.dme_estimate(cfg, h, s, m, (512, 1024), (1024, 1536), want_diag=True, form=form)
    blmin, blmax = 1.0 / 12.0, 36 * 4.0
    for l in range(L):
        est = np.array(diag[l].est_fp32)
        np.testing.assert_allclose(est[:, 0, :].sum(axis=0), 1.0, rtol=1e-5)
        vq = (h["bin_size"][l] ** 2) / 12.0
        assert np.all(est[0, 2, :] >= blmin) and np.all(est[0, 2, :] <= blmax + vq + 1e-3)
        # estimates accepted after at least one M-step have means s * relAmp
        after_m = (np.array(diag[l].converged) != 0) & (np.array(diag[l].iter_count) >= 1)
        assert after_m.sum() > 0
        ratio = est[1:, 1, after_m] / np.array([3.6, 2.4, 1.6, 1.0])[:, None]
        np.testing.assert_allclose(ratio, np.broadcast_to(ratio[0], ratio.shape), rtol=1e-6)
        conf = np.array(diag[l].confidence)
        assert np.all((conf == 0) | (conf >= cfg.success_confidence_thresh)) and np.all(conf <= 1)


# ---------------------------------------------------------------------------
# Synthetic generator -- the workload's stated distribution.
# ---------------------------------------------------------------------------

def test_generator_distribution(H, O, T):
    L, F = 8, 2048
    p = H.tracegen_default(drift_per_chunk=0.0)
    st = np.zeros(L, dtype=T.TRACEGEN_STATE)
    x, truth = O.tracegen(p, L, F, 0, st, True, truth=True)
    xb = x[truth == 0].astype(np.float64)
    assert abs(xb.mean()) < 0.3 and abs(xb.std() - 12) < 0.3
    frac = (truth > 0).mean()
    assert abs(frac - 8 / 28) < 0.02                   # geometric widths 8, gaps 20
    for a, amp in enumerate([180, 120, 80, 50]):
        xa = x[truth == a + 1].astype(np.float64)
        assert abs(xa.mean() - amp) < 0.02 * amp
        assert abs(np.sqrt(xa.var() - 144) - 0.1 * amp) < 0.02 * amp
        assert abs((truth == a + 1).mean() / frac - 0.25) < 0.02


def test_em_density_recurrence_error_bound():
    """The E-step evaluates each component density along the equally spaced bin
    boundaries by the exact identity exp(t(c+1)) = exp(t(c)) R(c), R(c+1) = R(c) Q
    (t quadratic in c), restarted from exp every 10 boundaries, with exp taken at every
    boundary of a block where the density rises out of the subnormal range.  Against
    float64 densities at the float32 boundaries the float32 products stay within 2e-4
    relative (the boundaries' own rounding, |x - mu| / var ulp(x), dominates), 1e-4
    within e^-20 of the peak, over component sigmas 5..40 and bin sizes 0.5..4."""
    f32 = np.float32
    rng = np.random.default_rng(11)
    nb, block = 300, 10
    worst = worst_peak = 0.0
    for _ in range(300):
        sigma = f32(rng.uniform(5.0, 40.0))
        bs = f32(rng.uniform(0.5, 4.0))
        lb = f32(rng.uniform(-60.0, 0.0))
        mu = f32(lb + rng.uniform(0.0, nb) * bs)
        var = f32(sigma * sigma)
        nhv = f32(f32(-0.5) * (f32(1.0) / var))
        nbs = f32(nhv * bs)
        pref = f32(-0.5 * np.log(var) - 0.9189385)
        x = (lb + np.arange(nb + 1, dtype=f32) * bs).astype(f32)
        y = (x - mu).astype(f32)
        t = (pref + nhv * (y * y)).astype(f32)
        q = f32(np.exp(f32(2.0) * nhv * (bs * bs)))
        exact = np.exp(np.float64(pref) - 0.5 * (x.astype(np.float64) - np.float64(mu)) ** 2 / np.float64(var))
        d = r = f32(0.0)
        ex = False
        for c in range(nb + 1):
            if c % block == 0:
                d = f32(np.exp(t[c]))
                dl = f32(nbs * f32(f32(2.0) * y[c] + bs))
                r = f32(np.exp(dl))
                ex = not ((d >= np.finfo(f32).tiny or dl <= 0) and dl <= 80)
            else:
                d = f32(np.exp(t[c])) if ex else f32(d * r)
                r = f32(r * q)
            if exact[c] > 1e-30:
                e = abs(float(d) / exact[c] - 1.0)
                worst = max(worst, e)
                if t[c] > pref - 20:
                    worst_peak = max(worst_peak, e)
    assert worst < 2e-4 and worst_peak < 1e-4, (worst, worst_peak)
