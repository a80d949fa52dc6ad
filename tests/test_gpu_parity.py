"""GPU parity: every stage through the HIP C-ABI against the CPU oracle on
the same seeded inputs.  Integer outputs and the full-precision model
outputs are compared bit for bit (the kernels keep the host operation
order and share the deterministic math of include/hwpa_math.h)."""
import numpy as np
import pytest

from conftest import gen_traces

pytestmark = pytest.mark.gpu


def _dev(H, a):
    return H.dev_from_numpy(a)


def _eq_struct(a, b, name):
    for f in a.dtype.names:
        x, y = a[f], b[f]
        if x.dtype.names:
            _eq_struct(x, y, name + "." + f)
            continue
        xb = np.ascontiguousarray(x).view(np.uint8)
        yb = np.ascontiguousarray(y).view(np.uint8)
        if not np.array_equal(xb, yb):
            bad = np.argwhere(np.ascontiguousarray(x) != np.ascontiguousarray(y))
            raise AssertionError("%s.%s differs at %d places, first %s: gpu=%r cpu=%r" % (
                name, f, len(bad), bad[:3].tolist(), x[tuple(bad[0])] if len(bad) else None,
                y[tuple(bad[0])] if len(bad) else None))


def test_tracegen_matches_host(H, O, T, gpu):
    L, F = 6, 512
    p = H.tracegen_default(seed=11, step_drift=60.0, step_span=1024.0, empty_fraction=0.05)
    st_h = np.zeros(L, dtype=T.TRACEGEN_STATE)
    d_st = H.dev_bytes(L * T.TRACEGEN_STATE.itemsize)
    d_out = H.dev_bytes(L * F * 64 * 2)
    for c in range(3):
        host = O.tracegen(p, L, F, c * F, st_h, c == 0)
        H.tracegen(p, L, F, c * F, d_st, c == 0, d_out)
        dev = H.to_numpy(d_out, np.int16, (L, F, 64))
        assert np.array_equal(dev, host), "chunk %d" % c


@pytest.mark.parametrize("ftype", [0, 1, 2, 3, 4, 5])
def test_baseliner_bit_exact(H, O, T, gpu, ftype):
    L, F, C = 4, 512, 4
    cfg = H.default_config(filter_type=ftype)
    chunks = gen_traces(O, T, H, L, F, C, seed=100 + ftype)
    ob = O.Baseliner(cfg, L)
    gb = H.Baseliner(cfg, L)
    d_out = H.dev_bytes(L * F * 64 * 2)
    d_stats = H.dev_bytes(L * T.BASELINER_STATS.itemsize)
    for c, tr in enumerate(chunks):
        o_out, o_stats = ob.run(tr, T.BASELINER_STATS)
        d_in = _dev(H, tr)
        gb.run(d_in, F, d_out, d_stats)
        g_out = H.to_numpy(d_out, np.int16, (L, F, 64))
        g_stats = H.to_numpy(d_stats, T.BASELINER_STATS, (L,))
        mism = np.argwhere(g_out != o_out)
        assert len(mism) == 0, "filter %d chunk %d: %d mismatches, first %s gpu=%d cpu=%d" % (
            ftype, c, len(mism), mism[0].tolist(), g_out[tuple(mism[0])], o_out[tuple(mism[0])])
        _eq_struct(g_stats, o_stats, "stats[c=%d]" % c)


@pytest.mark.parametrize("F", [512, 32, 16])
@pytest.mark.parametrize("variant", [1, 2, 3, 4, 5])
def test_baseliner_lane_tile_variants(H, O, T, gpu, variant, F):
    """2 and 4 lanes per workgroup (a lane count that leaves the last tile
    partial), one or two frame groups in flight, chunks of 1, 2 or 32 groups."""
    L, C = 5, 3
    H.tune_set(H.TUNE_BASELINE, L, F, variant)
    try:
        cfg = H.default_config()
        chunks = gen_traces(O, T, H, L, F, C, seed=77)
        ob, gb = O.Baseliner(cfg, L), H.Baseliner(cfg, L)
        d_out = H.dev_bytes(L * F * 64 * 2)
        d_stats = H.dev_bytes(L * T.BASELINER_STATS.itemsize)
        for tr in chunks:
            o_out, o_stats = ob.run(tr, T.BASELINER_STATS)
            gb.run(_dev(H, tr), F, d_out, d_stats)
            assert np.array_equal(H.to_numpy(d_out, np.int16, (L, F, 64)), o_out)
            _eq_struct(H.to_numpy(d_stats, T.BASELINER_STATS, (L,)), o_stats, "stats")
    finally:
        H.tune_set(H.TUNE_BASELINE, L, F, 0)


def test_autotune_picks_a_timed_variant(H, gpu):
    cfg = H.bench_config(512)
    r = H.autotune(cfg, 64, 512, repeats=2)
    counts = {"histogram": 22, "baseline": 6, "labels": 4, "pulses": 13}
    for i, (name, n) in enumerate(counts.items()):
        assert len(r[name]) == H.tune_variants(i) == n
        assert all(t > 0 for t in r[name])
        assert r[name][r["chosen"][i]] == min(r[name])
        H.tune_set(i, 64, 512, 0)


def test_baseliner_sigma_ema_and_scaler(H, O, T, gpu):
    L, F, C = 3, 256, 5
    cfg = H.default_config(mean_ema_scale_strides=4.0, jump_suppression=2.0)
    cfg.movie.photoelectron_sensitivity = 1.5
    cfg.movie.pedestal = 12
    chunks = gen_traces(O, T, H, L, F, C, seed=5)
    ob, gb = O.Baseliner(cfg, L), H.Baseliner(cfg, L)
    d_out = H.dev_bytes(L * F * 64 * 2)
    d_stats = H.dev_bytes(L * T.BASELINER_STATS.itemsize)
    for tr in chunks:
        o_out, o_stats = ob.run(tr, T.BASELINER_STATS)
        gb.run(_dev(H, tr), F, d_out, d_stats)
        assert np.array_equal(H.to_numpy(d_out, np.int16, (L, F, 64)), o_out)
        _eq_struct(H.to_numpy(d_stats, T.BASELINER_STATS, (L,)), o_stats, "stats")


def _pipeline_inputs(H, O, T, L, F, C, seed=3, cfg=None):
    """Baselined traces and chunk stats from the oracle baseliner."""
    cfg = cfg or H.bench_config(F)
    ob = O.Baseliner(cfg, L)
    out = []
    for tr in gen_traces(O, T, H, L, F, C, seed=seed):
        out.append(ob.run(tr, T.BASELINER_STATS))
    return cfg, out


def test_stats_aggregate_bit_exact(H, O, T, gpu):
    L, F, C = 5, 512, 5
    cfg, chunks = _pipeline_inputs(H, O, T, L, F, C)
    agg = O.Aggregator(L)
    d_agg = H.dev_bytes(L * T.BASELINER_STATS.itemsize)
    H.stats_reset(d_agg, L)
    for c, (_, st) in enumerate(chunks):
        agg.add(st)
        H.stats_merge(d_agg, _dev(H, st), L)
        _eq_struct(H.to_numpy(d_agg, T.BASELINER_STATS, (L,)), agg.get(T.BASELINER_STATS), "agg[%d]" % c)
        if c == 2:
            agg.reset()
            H.stats_reset(d_agg, L)


@pytest.fixture(params=range(22))
def hist_variant(request, H):
    """Every autotuner variant of the histogram kernel (shape 6 lanes x 512)."""
    H.tune_set(H.TUNE_HISTOGRAM, 6, 512, request.param)
    H.tune_set(H.TUNE_HISTOGRAM, 3, 512, request.param)
    yield request.param
    H.tune_set(H.TUNE_HISTOGRAM, 6, 512, 0)
    H.tune_set(H.TUNE_HISTOGRAM, 3, 512, 0)


def test_histogram_bit_exact(H, O, T, gpu, hist_variant):
    L, F, C = 6, 512, 5
    cfg, chunks = _pipeline_inputs(H, O, T, L, F, C, seed=21)
    hist = O.Histogram(cfg, L)
    d_hist = H.dev_bytes(L * T.LANE_HIST.itemsize)
    d_edge = H.dev_bytes(L * H.edge_state_bytes())
    H.edge_state_init(d_edge, L)
    agg = O.Aggregator(L)
    models = None
    for c, (bl, st) in enumerate(chunks):
        agg.add(st)
        s = agg.get(T.BASELINER_STATS)
        if models is None or c == 3:
            # reset from the running aggregate (bounds from baseline stats)
            assert hist.reset(s) == 0
            H.hist_reset(cfg, d_hist, _dev(H, s), L)
            models = O.dme_init_models(cfg, s, T.LANE_MODEL)
            continue
        hist.add(bl, models)
        H.hist_add(d_hist, d_edge, _dev(H, bl), F, _dev(H, models), L)
        _eq_struct(H.to_numpy(d_hist, T.LANE_HIST, (L,)), hist.get(T.LANE_HIST), "hist[%d]" % c)


@pytest.mark.parametrize("variant", [0, 9, 11, 12, 14, 16, 17, 19, 21])
@pytest.mark.parametrize("frames", [144, 48, 208])
def test_histogram_uneven_frames(H, O, T, gpu, variant, frames):
    """Chunk lengths that do not split evenly over a lane's waves (144 frames in 4 waves of
    16-frame batches: 48, 48, 48 and none; of 8-frame batches: 40, 40, 40, 24; 48 frames:
    too short for 4 waves of 16, one wave runs), over chunks with the edge-classifier
    state carried."""
    L, C = 3, 4
    H.tune_set(H.TUNE_HISTOGRAM, L, frames, variant)
    try:
        cfg, chunks = _pipeline_inputs(H, O, T, L, frames, C, seed=8)
        hist = O.Histogram(cfg, L)
        d_hist = H.dev_bytes(L * T.LANE_HIST.itemsize)
        d_edge = H.dev_bytes(L * H.edge_state_bytes())
        H.edge_state_init(d_edge, L)
        agg = O.Aggregator(L)
        for c, (bl, st) in enumerate(chunks):
            agg.add(st)
            s = agg.get(T.BASELINER_STATS)
            models = O.dme_init_models(cfg, s, T.LANE_MODEL)
            if c == 0:
                assert hist.reset(s) == 0
                H.hist_reset(cfg, d_hist, _dev(H, s), L)
            hist.add(bl, models)
            H.hist_add(d_hist, d_edge, _dev(H, bl), frames, _dev(H, models), L)
            _eq_struct(H.to_numpy(d_hist, T.LANE_HIST, (L,)), hist.get(T.LANE_HIST), "hist[%d]" % c)
        assert hist.get(T.LANE_HIST)["bin_count"].sum() > 0
    finally:
        H.tune_set(H.TUNE_HISTOGRAM, L, frames, 0)


def test_histogram_bin_edges_wrap_and_degenerate_bounds(H, O, T, gpu, hist_variant):
    """Samples exactly on and next to bin edges (1/binSize inexact), counters at
    the 16-bit wrap, and zero / NaN bin sizes, against the oracle."""
    L, F = 3, 512
    rng = np.random.default_rng(5)
    cfg = H.bench_config(F)
    h0 = np.zeros(L, dtype=T.LANE_HIST)
    h0["low_bound"] = -40.0
    h0["bin_size"] = rng.choice(np.array([2.4, 3.3, 1.7, 2.43, 0.3, 1.1, 7.9], dtype=np.float32), size=(L, 64))
    h0["bin_size"][2, :4] = [0.0, np.nan, 1e-30, 3e30]
    h0["low_bound"][2, 4:8] = [np.nan, np.inf, -1e9, 1e9]
    h0["bin_count"][1] = rng.integers(65000, 65536, size=(300, 64)).astype(np.uint16)   # wrap path
    h0["outlier_low"][1] = 65535
    h0["outlier_high"][1] = 65534
    # every integer in [-60, 450) plus the int16 extremes, shuffled per ZMW
    vals = np.concatenate([np.arange(-60, 450), [-32768, 32767]]).astype(np.int16)
    tr = np.stack([np.stack([rng.permutation(vals) for _ in range(64)], axis=1) for _ in range(L)])
    models = np.zeros(L, dtype=T.LANE_MODEL)
    models["baseline"]["means"] = np.float16(20000.0).view(np.uint16)   # every frame is "baseline": none are edges
    models["baseline"]["vars"] = np.float16(1.0).view(np.uint16)
    oh = O.Histogram(cfg, L)
    oh.set(h0)
    oh.add(tr, models)
    d_h = _dev(H, h0)
    d_edge = H.dev_bytes(L * H.edge_state_bytes())
    H.edge_state_init(d_edge, L)
    H.hist_add(d_h, d_edge, _dev(H, tr), F, _dev(H, models), L)
    got, want = H.to_numpy(d_h, T.LANE_HIST, (L,)), oh.get(T.LANE_HIST)
    assert want["bin_count"][0].sum() > 1000
    _eq_struct(got, want, "hist")


def test_dme_init_bit_exact(H, O, T, gpu):
    L, F, C = 4, 512, 3
    cfg, chunks = _pipeline_inputs(H, O, T, L, F, C, seed=9)
    agg = O.Aggregator(L)
    d_models = H.dev_bytes(L * T.LANE_MODEL.itemsize)
    for bl, st in chunks:
        agg.add(st)
        s = agg.get(T.BASELINER_STATS)
        H.dme_init_models(cfg, _dev(H, s), d_models, L)
        _eq_struct(H.to_numpy(d_models, T.LANE_MODEL, (L,)), O.dme_init_models(cfg, s, T.LANE_MODEL), "init")


def _em_case(H, O, T, L, F, seed, cfg=None, tg=None):
    cfg = cfg or H.bench_config(F)
    ob = O.Baseliner(cfg, L)
    hist = O.Histogram(cfg, L)
    agg = O.Aggregator(L)
    chunks = gen_traces(O, T, H, L, F, 3, seed=seed, **(tg or {}))
    bl0, st0 = ob.run(chunks[0], T.BASELINER_STATS)
    bl1, st1 = ob.run(chunks[1], T.BASELINER_STATS)
    agg.add(st1)
    s = agg.get(T.BASELINER_STATS)
    hist.reset(s)
    models = O.dme_init_models(cfg, s, T.LANE_MODEL)
    agg.reset()
    bl2, st2 = ob.run(chunks[2], T.BASELINER_STATS)
    agg.add(st2)
    hist.add(bl2, models)
    return cfg, hist.get(T.LANE_HIST), agg.get(T.BASELINER_STATS), models


@pytest.mark.parametrize("seed", [1, 2])
def test_dme_estimate_bit_exact(H, O, T, gpu, seed):
    L, F = 4, 512
    cfg, h, s, m = _em_case(H, O, T, L, F, seed)
    mfi, efi = (512, 1024), (1024, 1536)
    o_models, diag = O.dme_estimate(cfg, h, s, m, mfi, efi, want_diag=True, form="mirror")
    d_models = _dev(H, m)
    d_iters = H.dev_bytes(L * 4)
    H.dme_estimate(cfg, _dev(H, h), _dev(H, s), d_models, mfi, efi, L, d_iters)
    g_models = H.to_numpy(d_models, T.LANE_MODEL, (L,))
    iters = H.to_numpy(d_iters, np.int32, (L,))
    assert list(iters) == [diag[l].iterations_run for l in range(L)]
    _eq_struct(g_models, o_models, "models")


def test_dme_estimate_recurrence_edge_paths_bit_exact(H, O, T, gpu):
    """Histograms whose boundaries start far below every component (densities below
    FLT_MIN and rising at the first restarts of the E-step's density recurrence, so
    those blocks take exp at every boundary) and bins 8 or 16 times wider (ratios
    between boundaries beyond e^80 for the narrow components), against the oracle."""
    L, F = 4, 512
    cfg, h, s, m = _em_case(H, O, T, L, F, 3)
    h = h.copy()
    bs = h["bin_size"].copy()
    h["low_bound"][:, 0::2] -= 100.0 * bs[:, 0::2]
    h["bin_size"][1, 1::4] *= 8.0
    h["bin_size"][2, 1::4] *= 16.0
    h["low_bound"][3, 1::4] -= 250.0 * bs[3, 1::4]
    mfi, efi = (512, 1024), (1024, 1536)
    o_models, diag = O.dme_estimate(cfg, h, s, m, mfi, efi, want_diag=True, form="mirror")
    d_models = _dev(H, m)
    d_iters = H.dev_bytes(L * 4)
    H.dme_estimate(cfg, _dev(H, h), _dev(H, s), d_models, mfi, efi, L, d_iters)
    iters = H.to_numpy(d_iters, np.int32, (L,))
    assert list(iters) == [diag[l].iterations_run for l in range(L)]
    _eq_struct(H.to_numpy(d_models, T.LANE_MODEL, (L,)), o_models, "models")


@pytest.fixture(params=[0, 1, 2, 3])
def label_variant(request, H):
    """Every autotuner variant of the labeler (16 or 8 frames per batch, registers
    unconstrained or held to 4 waves per SIMD)."""
    for F in (512, 200):
        H.tune_set(H.TUNE_LABEL, 4, F, request.param)
    yield request.param
    for F in (512, 200):
        H.tune_set(H.TUNE_LABEL, 4, F, 0)


@pytest.mark.parametrize("F", [512, 200])
def test_labeler_bit_exact(H, O, T, gpu, label_variant, F):
    """Labels, latent frames and scores bit for bit; 200-frame chunks (the
    first 200 frames of baselined 208-frame chunks) end in partial batches."""
    L, C = 4, 4
    cfg, chunks = _pipeline_inputs(H, O, T, L, 512 if F == 512 else 208, C, seed=33)
    if F != 512:
        cfg = H.bench_config(F)
        chunks = [(np.ascontiguousarray(bl[:, :F, :]), st) for bl, st in chunks]
    agg = O.Aggregator(L)
    ol = O.Labeler(cfg, L)
    gl = H.Labeler(cfg, L, F)
    d_lab = H.dev_bytes(L * F * 64 * 2)
    d_lat = H.dev_bytes(L * 24 * 64 * 2)
    d_score = H.dev_bytes(L * 64 * 4)
    for c, (bl, st) in enumerate(chunks):
        agg.add(st)
        models = O.dme_init_models(cfg, agg.get(T.BASELINER_STATS), T.LANE_MODEL)
        o_lab, o_lat, o_score = ol.run(bl, models)
        gl.run(_dev(H, bl), F, _dev(H, models), d_lab, d_lat, d_score)
        g_lab = H.to_numpy(d_lab, np.int16, (L, F, 64))
        mism = np.argwhere(g_lab != o_lab)
        assert len(mism) == 0, "chunk %d: %d label mismatches, first %s" % (c, len(mism), mism[:3].tolist())
        assert np.array_equal(H.to_numpy(d_lat, np.int16, (L, 24, 64)), o_lat)
        assert np.array_equal(H.to_numpy(d_score, np.float32, (L, 64)).view(np.uint32), o_score.view(np.uint32))


@pytest.mark.parametrize("variant", [0, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12])
@pytest.mark.parametrize("max_calls", [0, 480])
def test_pulses_dense_boundaries_bit_exact(H, O, T, gpu, variant, max_calls):
    """A pulse boundary at every frame (up states cycling through the analogs, with a
    baseline frame every 7th): about one record per frame and ZMW, so the queued variants
    fill their LDS queues to capacity before flushes; with the default capacity every ZMW
    overflows, with 480 slots every record is kept."""
    L, F, C = 2, 512, 2
    H.tune_set(H.TUNE_PULSE, L, F, variant)
    try:
        cfg = H.bench_config(F)
        if max_calls:
            cfg.max_calls_per_zmw = max_calls
        cfg, chunks = _pipeline_inputs(H, O, T, L, F, C, seed=45, cfg=cfg)
        agg = O.Aggregator(L)
        op, gp = O.Pulser(cfg, L), H.Pulser(cfg, L)
        n, mc = L * 64, cfg.max_calls_per_zmw
        d_p = H.dev_bytes(n * mc * T.PULSE.itemsize)
        d_c, d_o = H.dev_bytes(n * 4), H.dev_bytes(n * 4)
        d_bl = H.dev_bytes(L * T.STAT_ACC.itemsize)
        f = np.arange(F)[None, :, None] + np.arange(64)[None, None, :]
        lab = np.where(f % 7 == 6, 0, 5 + f % 4).astype(np.int16) * np.ones((L, 1, 1), np.int16)
        lat = np.zeros((L, 24, 64), np.int16)
        for c, (bl, st) in enumerate(chunks):
            agg.add(st)
            models = O.dme_init_models(cfg, agg.get(T.BASELINER_STATS), T.LANE_MODEL)
            ff = c * F - 24
            o_p, o_c, o_o, o_bl = op.run(lab, lat, bl, ff, models, T.PULSE, T.STAT_ACC)
            gp.run(_dev(H, lab), _dev(H, lat), _dev(H, bl), F, ff, _dev(H, models), d_p, d_c, d_o, d_bl)
            assert np.array_equal(H.to_numpy(d_c, np.uint32, (n,)), o_c)
            assert np.array_equal(H.to_numpy(d_o, np.uint32, (n,)), o_o)
            assert (o_c + o_o).min() >= F * 6 // 7 - 2
            assert (o_o.min() > 0) if not max_calls else (o_o.max() == 0)
            g_p = H.to_numpy(d_p, T.PULSE, (n, mc))
            for z in range(n):
                k = o_c[z]
                assert np.array_equal(np.ascontiguousarray(g_p[z, :k]).view(np.uint8),
                                      np.ascontiguousarray(o_p[z, :k]).view(np.uint8)), (c, z)
            _eq_struct(H.to_numpy(d_bl, T.STAT_ACC, (L,)), o_bl, "pulse baseline stats")
    finally:
        H.tune_set(H.TUNE_PULSE, L, F, 0)


@pytest.mark.parametrize("variant,max_calls", [(0, 0), (1, 0), (2, 0), (3, 0), (4, 0), (5, 0), (6, 0), (7, 0), (8, 0), (9, 0),
                                              (10, 0), (11, 0), (12, 0), (0, 5), (3, 5), (4, 5), (5, 5), (7, 5), (9, 5),
                                              (10, 5), (12, 5)])
def test_pulses_bit_exact(H, O, T, gpu, variant, max_calls):
    """Pulse records, counts, overflow counts and baseline stats of every pulse-kernel
    variant; max_calls 5 makes most ZMWs overflow their record capacity."""
    L, F, C = 4, 512, 4
    H.tune_set(H.TUNE_PULSE, L, F, variant)
    cfg = H.bench_config(F)
    if max_calls:
        cfg.max_calls_per_zmw = max_calls
    cfg, chunks = _pipeline_inputs(H, O, T, L, F, C, seed=44, cfg=cfg)
    agg = O.Aggregator(L)
    ol = O.Labeler(cfg, L)
    op = O.Pulser(cfg, L)
    gp = H.Pulser(cfg, L)
    n = L * 64
    mc = cfg.max_calls_per_zmw
    d_p = H.dev_bytes(n * mc * T.PULSE.itemsize)
    d_c, d_o = H.dev_bytes(n * 4), H.dev_bytes(n * 4)
    d_bl = H.dev_bytes(L * T.STAT_ACC.itemsize)
    for c, (bl, st) in enumerate(chunks):
        agg.add(st)
        models = O.dme_init_models(cfg, agg.get(T.BASELINER_STATS), T.LANE_MODEL)
        lab, lat, _ = ol.run(bl, models)
        ff = c * F - 24
        o_p, o_c, o_o, o_bl = op.run(lab, lat, bl, ff, models, T.PULSE, T.STAT_ACC)
        gp.run(_dev(H, lab), _dev(H, lat), _dev(H, bl), F, ff, _dev(H, models), d_p, d_c, d_o, d_bl)
        g_c = H.to_numpy(d_c, np.uint32, (n,))
        assert np.array_equal(g_c, o_c), "chunk %d counts" % c
        assert np.array_equal(H.to_numpy(d_o, np.uint32, (n,)), o_o)
        if max_calls:
            assert o_o.sum() > 0
        g_p = H.to_numpy(d_p, T.PULSE, (n, mc))
        for z in range(n):
            k = o_c[z]
            if k:
                a = np.ascontiguousarray(g_p[z, :k]).view(np.uint8)
                b = np.ascontiguousarray(o_p[z, :k]).view(np.uint8)
                assert np.array_equal(a, b), "chunk %d zmw %d pulses differ" % (c, z)
        _eq_struct(H.to_numpy(d_bl, T.STAT_ACC, (L,)), o_bl, "pulse baseline stats")
    H.tune_set(H.TUNE_PULSE, L, F, 0)


@pytest.mark.parametrize("hist_v", [0, 16, 19])
def test_analyzer_end_to_end(H, O, T, gpu, hist_v):
    """Whole analyzer (bench configuration: one estimate per chunk) against
    the CPU analyzer, every chunk, every output; with the one-wave histogram and
    with the all-bins-in-LDS forms, whose counters start from the reset that
    leaves the zeros to them."""
    L, F, C = 8, 512, 7
    H.tune_set(H.TUNE_HISTOGRAM, L, F, hist_v)
    try:
        _analyzer_end_to_end(H, O, T, L, F, C)
    finally:
        H.tune_set(H.TUNE_HISTOGRAM, L, F, 0)


def _analyzer_end_to_end(H, O, T, L, F, C):
    cfg = H.bench_config(F)
    chunks = gen_traces(O, T, H, L, F, C, seed=77)
    oa = O.Analyzer(cfg, 0, L, F, em="mirror")
    ga = H.BatchAnalyzer(0, L, F, cfg)
    dts = (T.PULSE, T.STAT_ACC, T.BASELINER_STATS, T.LANE_MODEL)
    for c, tr in enumerate(chunks):
        r = oa(tr, c * F, dts, want=("pulses", "labels", "models", "baselined", "score"))
        ga(tr, c * F)
        assert ga.flags()[:3] == r["flags"][:3], "chunk %d flags" % c
        assert np.array_equal(ga.baselined(), r["baselined"]), "chunk %d baselined" % c
        if r["flags"][2]:
            _eq_struct(ga.models(), r["models"], "models[%d]" % c)
            lab = ga.labels()
            assert np.array_equal(lab, r["labels"]), "chunk %d labels: %d differ" % (c, int((lab != r["labels"]).sum()))
        p, cnt, ovf = ga.pulses()
        assert np.array_equal(cnt, r["counts"]), "chunk %d pulse counts" % c
        for z in np.nonzero(cnt)[0]:
            k = cnt[z]
            assert np.array_equal(np.ascontiguousarray(p[z, :k]).view(np.uint8),
                                  np.ascontiguousarray(r["pulses"][z, :k]).view(np.uint8)), "chunk %d zmw %d" % (c, z)


def test_analyzer_rejects_bad_pool_and_frame(H, T, gpu):
    cfg = H.bench_config(512)
    ga = H.BatchAnalyzer(3, 2, 512, cfg)
    tr = np.zeros((2, 512, 64), dtype=np.int16)
    with pytest.raises(H.HwpaError) as e:
        ga(tr, 0, pool_id=4)
    assert e.value.code == T.ERR_BAD_POOL
    with pytest.raises(H.HwpaError) as e:
        ga(tr, 512)
    assert e.value.code == T.ERR_BAD_FRAME
    ga(tr, 0)
    ga(tr, 512)
