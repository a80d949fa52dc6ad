"""Known answers and closed forms for oracle functions the pipeline tests cannot
pin on their own: EvolveConfidence, the model updates, the confidence factors,
the histogram fractile / cumulative count and the preliminary scale factor,
the subframe scorer, the ROI hysteresis and the pool stagger.  Expected values
come from the reference's own unit tests where it has one, else from the
definition worked out here in float64 (or by hand), with inputs chosen so that
a swapped index, a dropped term or a wrong branch changes the result."""
import ctypes

import numpy as np
import pytest

FORMS = ("reference", "mirror")


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _m16(bl, an, conf):
    """bl = (w, mean, var); an = 4 x (w, mean, var) -> the oracle's 16-float model."""
    return np.array(list(bl) + [v for a in an for v in a] + [conf], dtype=np.float32)


def _update(O, form, cfg, t, o, fraction):
    out = np.zeros(16, dtype=np.float32)
    O.lib.orc_dme_update(ctypes.c_int32(O.FORMS[form]), ctypes.byref(cfg), _p(t), _p(o),
                         ctypes.c_float(fraction), _p(out))
    return out


# ---------------------------------------------------------------------------
# EvolveConfidence (basecaller/test/TestDetectionModelHost.cpp:118-127)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("form", FORMS)
def test_evolve_confidence_reference_known_answer(H, O, form):
    O.lib.orc_dme_evolve.restype = ctypes.c_float
    cfg = H.default_config()

    def evolve(mfi, efi, hl):
        cfg.confidence_half_life[0] = cfg.confidence_half_life[1] = hl      # hl0 == hl1: no ramp
        return O.lib.orc_dme_evolve(ctypes.c_int32(O.FORMS[form]), ctypes.byref(cfg),
                                    (ctypes.c_int32 * 2)(*mfi), (ctypes.c_int32 * 2)(*efi))

    conf = 1.0 * evolve((0, 100), (100, 200), 100.0)
    assert conf == 0.5
    conf *= evolve((100, 200), (200, 300), 200.0)
    expect = 0.5 * np.sqrt(0.5)
    assert abs(conf - expect) < 1e-6 * expect


@pytest.mark.parametrize("form", FORMS)
def test_evolve_half_life_ramp(H, O, form):
    """hl = hl0 + (hl1 - hl0) satlin(56160, 168480, t) at the midpoint t of the two
    interval centres (DmeEmHost.cpp:806-834): hl0 below, hl1 above, their mean at
    the ramp's middle."""
    O.lib.orc_dme_evolve.restype = ctypes.c_float
    cfg = H.default_config()
    hl0, hl1 = cfg.confidence_half_life[0], cfg.confidence_half_life[1]
    for mfi, efi, hl in [((0, 2), (2000, 2002), hl0),                      # t ~ 1000
                         ((300000, 300002), (302000, 302002), hl1),        # t ~ 301000
                         ((0, 2), (224638, 224640), 0.5 * (hl0 + hl1))]:   # t = 112320
        tm = mfi[0] + (mfi[1] - mfi[0]) // 2
        te = efi[0] + (efi[1] - efi[0]) // 2
        got = O.lib.orc_dme_evolve(ctypes.c_int32(O.FORMS[form]), ctypes.byref(cfg),
                                   (ctypes.c_int32 * 2)(*mfi), (ctypes.c_int32 * 2)(*efi))
        np.testing.assert_allclose(got, 2.0 ** (-abs(te - tm) / hl), rtol=2e-6)


# ---------------------------------------------------------------------------
# DetectionModelHost::Update0/1/2 and Update(other) (DetectionModelHost.cpp:69-223)
# ---------------------------------------------------------------------------

T0 = _m16((0.6, 1.5, 100.0), [(0.1, 70.0, 260.0), (0.1, 50.0, 190.0), (0.1, 33.0, 150.0), (0.1, 21.0, 125.0)], 0.3)
O0 = _m16((0.5, -2.0, 140.0), [(0.12, 80.0, 320.0), (0.12, 55.0, 230.0), (0.13, 36.0, 190.0), (0.13, 24.0, 168.0)],
          0.9)


def _covar(mean, blvar, cv2, shot=1.0):
    return blvar + shot * mean + cv2 * mean * mean


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("frac", [0.0, 1.0, 0.5, 0.25])
def test_update_geometric_closed_form(H, O, form, frac):
    """Method 1 (the default): baseline weight and mean blended linearly, baseline
    variance and analog means geometrically, analog weights (1 - w_bl) / 4,
    analog variances from ModelSignalCovar; fraction 0 keeps the target, 1
    takes the other exactly."""
    cfg = H.default_config()
    assert cfg.model_update_method == 1
    got = _update(O, form, cfg, T0, O0, frac).astype(np.float64)
    t, o = T0.astype(np.float64), O0.astype(np.float64)
    a, b = 1.0 - frac, frac
    bw = a * t[0] + b * o[0]
    bv = t[2] ** a * o[2] ** b
    exp = np.zeros(16)
    exp[0:3] = [bw, a * t[1] + b * o[1], bv]
    for i in range(4):
        am = t[4 + 3 * i] ** a * o[4 + 3 * i] ** b
        exp[3 + 3 * i: 6 + 3 * i] = [0.25 * (1 - bw), am, _covar(am, bv, 0.01)]
    exp[15] = t[15]                               # Update(other, fraction) leaves the confidence
    np.testing.assert_allclose(got, exp, rtol=2e-6)
    if frac == 0.0:
        assert np.array_equal(got[[1, 2, 4, 7, 10, 13]], T0[[1, 2, 4, 7, 10, 13]])
    if frac == 1.0:
        assert np.array_equal(got[[1, 2, 4, 7, 10, 13]], O0[[1, 2, 4, 7, 10, 13]])


@pytest.mark.parametrize("form", FORMS)
def test_update_arithmetic_and_xsn_closed_forms(H, O, form):
    """Method 0 blends means and variances linearly and leaves the weights;
    method 2 blends the excess-noise CV^2 implied by each model
    ((var - blVar - mean * shot) / mean^2, the target's against its own
    pre-update baseline variance) and rebuilds the analog variances from it."""
    frac = 0.4
    a, b = 1 - frac, frac
    t, o = T0.astype(np.float64), O0.astype(np.float64)
    cfg = H.default_config()
    cfg.model_update_method = 0
    got = _update(O, form, cfg, T0, O0, frac).astype(np.float64)
    exp = t.copy()
    for k in [1, 2] + [4 + 3 * i for i in range(4)] + [5 + 3 * i for i in range(4)]:
        exp[k] = a * t[k] + b * o[k]
    np.testing.assert_allclose(got, exp, rtol=2e-6)
    cfg.model_update_method = 2
    got = _update(O, form, cfg, T0, O0, frac).astype(np.float64)
    bw = a * t[0] + b * o[0]
    bv = t[2] ** a * o[2] ** b
    for i in range(4):
        tm, tv, om, ov = t[4 + 3 * i], t[5 + 3 * i], o[4 + 3 * i], o[5 + 3 * i]
        cv2 = a * (tv - t[2] - tm) / tm ** 2 + b * (ov - o[2] - om) / om ** 2
        am = tm ** a * om ** b
        np.testing.assert_allclose(got[3 + 3 * i: 6 + 3 * i], [0.25 * (1 - bw), am, _covar(am, bv, cv2)], rtol=2e-5)


@pytest.mark.parametrize("form", FORMS)
def test_update_confidence_weighted(H, O, form):
    """Update(other): fraction = c_o / (c_t + c_o), confidences add.  With two
    zero confidences the fraction is 0: the header says the target is then not
    modified, but the code still applies Update(other, 0), which keeps the
    baseline and every analog mean and rebuilds the analog weights and
    variances by the model rule (DetectionModelHost.cpp:203-223, 98-134)."""
    cfg = H.default_config()
    got = _update(O, form, cfg, T0, O0, -1.0)
    frac = np.float32(O0[15]) / (np.float32(T0[15]) + np.float32(O0[15]))
    want = _update(O, form, cfg, T0, O0, float(frac))
    assert np.array_equal(got[:15], want[:15])
    assert got[15] == np.float32(T0[15]) + np.float32(O0[15])
    t, o = T0.copy(), O0.copy()
    t[15] = o[15] = 0.0
    got = _update(O, form, cfg, t, o, -1.0)
    assert np.array_equal(got, _update(O, form, cfg, t, o, 0.0))
    assert np.array_equal(got[[0, 1, 2, 4, 7, 10, 13, 15]], t[[0, 1, 2, 4, 7, 10, 13, 15]])
    np.testing.assert_allclose(got[[5, 8, 11, 14]], [_covar(t[k], t[2], 0.01) for k in (4, 7, 10, 13)], rtol=1e-6)


# ---------------------------------------------------------------------------
# DmeEmHost::ComputeConfidence (DmeEmHost.cpp:717-803)
# ---------------------------------------------------------------------------

def _satlin(a, b, x):
    return min(max((x - a) / (b - a), 0.0), 1.0)


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("converged", [0, 1])
def test_confidence_factors(H, O, form, converged):
    cfg = H.default_config()
    ref = _m16((0.6, 0.5, 10.0), [(0.1, 40.0, 60.0), (0.1, 30.0, 50.0), (0.1, 20.0, 40.0), (0.1, 12.0, 30.0)], 0.7)
    est = _m16((0.08, 2.0, 16.0), [(0.02, 32.0, 60.0), (0.05, 25.0, 50.0), (0.04, 18.0, 40.0), (0.03, 12.0, 30.0)], 0.0)
    cf = np.zeros(8, dtype=np.float32)
    O.lib.orc_dme_confidence(ctypes.c_int32(O.FORMS[form]), ctypes.byref(cfg), ctypes.c_int32(converged), _p(ref),
                             _p(est), _p(cf))
    r, e = ref.astype(np.float64), est.astype(np.float64)
    m0, m1 = cfg.analog_mix_fraction_threshold[0], cfg.analog_mix_fraction_threshold[1]
    sigma, ref_sigma = np.sqrt(e[2]), np.sqrt(r[2])
    tol = 1.5 / (0.5 + r[15])
    ref_snr0 = r[4] / ref_sigma
    ref_snr1 = r[7] / ref_sigma * cfg.snr_drop_thresh * min(r[15], 1.0)
    exp = [float(converged),
           _satlin(0.05, 0.15, e[0]),
           np.exp(-(e[1] ** 2 / e[2]) / 2.0),
           np.exp(-np.log2(e[2] / r[2]) ** 2 / (2 * tol * tol)),
           _satlin(m0, m1, e[3]) * _satlin(m0, m1, e[12]),
           _satlin(cfg.min_analog_snr_thresh0, cfg.min_analog_snr_thresh1, e[13] / sigma),
           _satlin(ref_snr1, np.sqrt(ref_snr0 * ref_snr1), e[4] / sigma),
           1.0]
    assert all(0.0 < x < 1.0 for x in exp[1:7])           # no factor saturated: every input matters
    np.testing.assert_allclose(cf, exp, rtol=3e-6)


# ---------------------------------------------------------------------------
# UHistogramSimd::Fractile / CumulativeCount (UHistogramSimd.cpp:140-242) and
# DmeEmHost::PrelimScaleFactor (DmeEmHost.cpp:618-654)
# ---------------------------------------------------------------------------

def _two_bin_hist(T, lo, bs, bins, counts, low=0, high=0):
    h = np.zeros(1, dtype=T.LANE_HIST)
    h["low_bound"][0, :], h["bin_size"][0, :] = lo, bs
    for j, n in zip(bins, counts):
        h["bin_count"][0, j, :] = n
    h["outlier_low"][0, :], h["outlier_high"][0, :] = low, high
    return h


def test_fractile_closed_form(O, T):
    """One populated bin j with N counts and K low outliers: the f-fractile is
    start_j + bs (f total - K) / (N + 1) while K < f total <= K + N; at or below
    the low outliers -inf, beyond the in-range counts +inf."""
    lo, bs, j, N, K, Hh = -10.0, 0.5, 37, 200, 30, 20
    h = _two_bin_hist(T, lo, bs, [j], [N], low=K, high=Hh)
    total = N + K + Hh
    frac = np.array([0.2, 0.5, 0.75] + [0.05, 0.99] * 30 + [0.3], dtype=np.float32)
    out = np.zeros(64, dtype=np.float32)
    O.lib.orc_hist_fractile(_p(h), _p(frac), _p(out))
    for f, g in zip(frac, out):
        nfz = np.float32(f) * np.float32(total)
        if nfz <= K:
            assert g == -np.inf
        elif nfz > K + N:
            assert g == np.inf
        else:
            np.testing.assert_allclose(g, lo + bs * j + bs * (nfz - K) / (N + 1), rtol=1e-6)


def test_cumulative_count_closed_form(O, T):
    lo, bs, j, N, K = -10.0, 0.5, 37, 200, 30
    h = _two_bin_hist(T, lo, bs, [j], [N], low=K, high=5)
    x0 = lo + bs * j
    xs = np.array([x0 + 0.25 * bs, x0, x0 + 0.9 * bs, lo - 1.0, lo + 300 * bs + 1.0, x0 - bs, np.nan]
                  + [x0 + 0.5 * bs] * 57, dtype=np.float32)
    out = np.zeros(64, dtype=np.float32)
    O.lib.orc_hist_cumulative(_p(h), _p(xs), _p(out))
    np.testing.assert_allclose(out[[0, 1, 2, 7]], [K + N * 0.25, K, K + N * 0.9, K + N * 0.5], rtol=1e-6)
    assert out[3] == K and out[4] == K + N and out[5] == K and np.isnan(out[6])


@pytest.mark.parametrize("conf", [0.0, 0.5])
def test_prelim_scale_factor_closed_form(H, O, T, conf):
    """Baseline counts Nb in a bin below mu_B + 2.5 sigma, pulse counts Np in a
    bin above it: the fractile 0.5 (Nb + total) / total lands in the pulse bin at
    start_p + bs (Np / 2) / (Np + 1); the factor is that over the mean analog
    mean, blended toward 1 by satlin(0, ScaleSnrConfTol, confidence) and
    clamped to [0.1, 10]."""
    cfg = H.default_config()
    lo, bs, jb, jp, Nb, Np = -20.0, 1.0, 20, 120, 400, 100
    h = _two_bin_hist(T, lo, bs, [jb, jp], [Nb, Np])
    means = np.array([90.0, 60.0, 40.0, 25.0])
    models = np.tile(_m16((0.6, 0.0, 16.0), [(0.1, m, 100.0) for m in means], conf), 64)
    models[16 * 5 + 4] = 0.1                    # ZMW 5: tiny analog means -> clamp at 10
    models[16 * 5 + 7] = models[16 * 5 + 10] = models[16 * 5 + 13] = 0.1
    models[16 * 6 + 4] = models[16 * 6 + 7] = models[16 * 6 + 10] = models[16 * 6 + 13] = 1e6   # -> 0.1
    out = np.zeros(64, dtype=np.float32)
    O.lib.orc_prelim_scale_factor(ctypes.byref(cfg), _p(h), _p(models), _p(out))
    frac_x = lo + bs * jp + bs * (Np / 2.0) / (Np + 1)
    w = _satlin(0.0, cfg.scale_snr_conf_tol, conf)
    sf = (1 - w) * frac_x / means.mean() + w
    np.testing.assert_allclose(out[0], min(max(sf, 0.1), 10.0), rtol=2e-6)
    assert out[5] == 10.0
    if conf == 0.0:
        assert out[6] == np.float32(0.1)
    else:
        np.testing.assert_allclose(out[6], (1 - w) * frac_x / 1e6 + w, rtol=2e-6)


# ---------------------------------------------------------------------------
# BlockStateSubframeScorer (SubframeScorer.h:143-318)
# ---------------------------------------------------------------------------

def _model(T, bl, an):
    m = np.zeros(1, dtype=T.LANE_MODEL)
    m["baseline"]["means"][0, :] = np.float16(bl[0]).view(np.uint16)
    m["baseline"]["vars"][0, :] = np.float16(bl[1]).view(np.uint16)
    for a, (mu, v) in enumerate(an):
        m["analogs"]["means"][0, a, :] = np.float16(mu).view(np.uint16)
        m["analogs"]["vars"][0, a, :] = np.float16(v).view(np.uint16)
    return m


def _normal_log(x, mu, v):
    return -0.5 * np.log(2 * np.pi * v) - 0.5 * (x - mu) ** 2 / v


@pytest.mark.parametrize("dim", [False, True])
def test_scorer_closed_forms(O, T, dim):
    """Baseline and full-frame states are the normal log density; an up or a
    down state of analog a (states 5+a and 9+a score alike) is the Gauss-cap
    subframe score: -log(mu_a - mu_B) between the joints x0 = mu_B + sigma_B
    sqrt(pi/2) and x1 = mu_a - sigma_a sqrt(pi/2), with the baseline (below x0) or
    analog (above x1) quadratic outside, capped at min(., full-frame - 1, 0) above
    mu_a.  With x1 < x0 for the dimmest analog (dim) its score falls back to the
    normal of mean (mu_B + mu_3)/2 and variance ((mu_3 - mu_B)/6 + v_B + v_3)/2."""
    an = [(180.0, 520.0), (120.0, 370.0), (80.0, 260.0), (14.0 if dim else 50.0, 196.0 if dim else 200.0)]
    bl = (1.0, 144.0)
    m = _model(T, bl, an)
    x = np.linspace(-60.0, 260.0, 1601).astype(np.float32)
    out = np.zeros((len(x), 13), dtype=np.float32)
    O.lib.orc_scorer_scores(_p(m), ctypes.c_int32(7), _p(x), ctypes.c_int32(len(x)), _p(out))
    xd = x.astype(np.float64)
    np.testing.assert_allclose(out[:, 0], _normal_log(xd, *bl), rtol=1e-5, atol=1e-4)
    x0 = bl[0] + np.sqrt(bl[1]) * np.sqrt(np.pi / 2)
    for a, (mu, v) in enumerate(an):
        full = _normal_log(xd, mu, v)
        np.testing.assert_allclose(out[:, 1 + a], full, rtol=1e-5, atol=1e-4)
        assert np.array_equal(out[:, 5 + a], out[:, 9 + a])
        x1 = mu - np.sqrt(v) * np.sqrt(np.pi / 2)
        s = np.where(xd < x0, -0.5 * (xd - x0) ** 2 / bl[1],
                     np.where(xd > x1, -0.5 * (xd - x1) ** 2 / v, 0.0)) - np.log(mu - bl[0])
        if a == 3 and x1 < x0:
            em, ev = 0.5 * (bl[0] + mu), 0.5 * ((mu - bl[0]) / 6.0 + bl[1] + v)
            s = _normal_log(xd, em, ev)
        s = np.where(xd > mu, np.minimum(np.minimum(s, full - 1.0), 0.0), s)
        np.testing.assert_allclose(out[:, 5 + a], s, rtol=1e-5, atol=2e-4)
    assert dim == (an[3][0] - np.sqrt(an[3][1]) * np.sqrt(np.pi / 2) < x0)


# ---------------------------------------------------------------------------
# ROI hysteresis (RoiHost.h:137-248), thresholds upper 7 / lower 2
# ---------------------------------------------------------------------------

def _roi(O, seq, before=0.0, boundary=0, upper=7.0, lower=2.0):
    seq = np.ascontiguousarray(seq, dtype=np.float32)
    n = len(seq) - 2
    out = np.zeros(n, dtype=np.uint8)
    O.lib.orc_roi(ctypes.c_float(upper), ctypes.c_float(lower), ctypes.c_float(before), ctypes.c_uint16(boundary),
                  _p(seq), ctypes.c_int32(n), _p(out))
    return out


def _frames(out):
    return set(np.nonzero(out)[0].tolist())


def test_roi_step_worked_by_hand(O):
    """A step of 5 sigma on frames [20, 40): inside, the 4-frame window gives
    2 x 5 = 10 >= 7; frame 19 sees (x18..x21)/2 = 5 and frame 18 (x17..x20)/2 = 2.5,
    both >= 2 and joined to the step by the backward pass; frame 40 sees
    (x39 + x40)/sqrt 2 = 3.5 >= 2 after a strong frame (hysteresis); frame 41's
    windows hold no step frame.  ROI = [18, 41)."""
    x = np.zeros(82)
    x[20:40] = 5.0
    assert _frames(_roi(O, x)) == set(range(18, 41))


def test_roi_needs_the_upper_threshold(O):
    """3 sigma (window value 6 < 7) never opens an ROI; 3.5 sigma (7) does."""
    x = np.zeros(82)
    x[20:40] = 3.0
    assert _frames(_roi(O, x)) == set()
    x[20:40] = 3.5
    assert 20 in _frames(_roi(O, x))


def test_roi_hysteresis_forward_and_backward(O):
    """A 1.5 sigma plateau (pair window 2.12 >= 2, below 7) stays in the ROI after
    a strong step (forward hysteresis) and before one (backward pass), and is
    out of it alone."""
    x = np.zeros(102)
    x[20:30] = 5.0
    x[30:60] = 1.5
    r = _frames(_roi(O, x))
    assert set(range(18, 60)) <= r and not (r & set(range(0, 17))) and not (r & set(range(62, 100)))
    x = np.zeros(102)
    x[10:40] = 1.5
    x[40:50] = 5.0
    r = _frames(_roi(O, x))
    assert set(range(10, 50)) <= r and not (r & set(range(0, 8)))
    x = np.zeros(102)
    x[10:40] = 1.5
    assert _frames(_roi(O, x)) == set()


def test_roi_boundary_carries_the_state(O):
    """A plateau at the start of a chunk continues an ROI the previous chunk left
    open (boundary bit set), and opens none otherwise."""
    x = np.zeros(62)
    x[0:30] = 1.5
    assert _frames(_roi(O, x, before=1.5, boundary=1)) >= set(range(0, 29))
    assert _frames(_roi(O, x, before=1.5, boundary=0)) == set()


# ---------------------------------------------------------------------------
# DynamicEstimateBatchAnalyzer stagger (BatchAnalyzer.cpp:114-117) and the
# estimator's frame counting (DetectionModelEstimator.cpp:93-160)
# ---------------------------------------------------------------------------

def test_pool_stagger_first_estimate_chunk(H, O, T):
    """Sigma EMA off (baseliner start-up 1024 frames), 1000 pre-accumulation and
    4000 estimation frames, 512-frame chunks: ceil(4000/512) = 8 chunks per
    estimate, so pool p waits (p mod 8) chunks more.  It starts feeding the
    estimator at chunk 2 + (p mod 8); 2 chunks cover the 1000 pre-accumulation
    frames and 8 the 4000 estimation frames, so the first estimate runs at chunk
    11 + (p mod 8)."""
    cfg = H.default_config()
    cfg.sigma_ema_scale_strides = 0.0
    L, F = 1, 512
    p = H.tracegen_default(seed=3)
    dts = (T.PULSE, T.STAT_ACC, T.BASELINER_STATS, T.LANE_MODEL)
    chunks = []
    st = np.zeros(L, dtype=T.TRACEGEN_STATE)
    for c in range(20):
        chunks.append(O.tracegen(p, L, F, c * F, st, c == 0))
    for pool in (0, 1, 3, 7, 8, 9):
        an = O.Analyzer(cfg, pool, L, F)
        first = None
        for c, tr in enumerate(chunks):
            r = an(tr, c * F, dts, want=())
            if r["flags"][1] and first is None:
                first = c
        assert first == 11 + pool % 8, (pool, first)
