"""GPU parity of the HF-metrics stage (hwpa_metrics_*) and of the analyzer's
metrics output against the CPU oracle: every field bit for bit, except that
NaN equals NaN (x86 and CDNA produce NaNs of different sign)."""
import numpy as np
import pytest

import metrics_inputs as MI
from conftest import gen_traces

pytestmark = pytest.mark.gpu


def _eq_metrics(g, o, name):
    for f in g.dtype.names:
        x, y = np.ascontiguousarray(g[f]), np.ascontiguousarray(o[f])
        if x.dtype.kind == "f":
            same = (x.view(np.uint32) == y.view(np.uint32)) | (np.isnan(x) & np.isnan(y))
        else:
            same = x == y
        if not same.all():
            bad = np.argwhere(~same)
            raise AssertionError("%s.%s differs at %d places, first %s: gpu=%r cpu=%r" % (
                name, f, len(bad), bad[:3].tolist(), x[tuple(bad[0])], y[tuple(bad[0])]))


def _tree(H):
    """A hand-made depth-3 activity tree over several features."""
    return H.activity_tree(children_left=[1, 3, 5, -1, -1, 7, -1, -1, -1],
                           children_right=[2, 4, 6, -1, -1, 8, -1, -1, -1],
                           feature=[0, 4, 8, -2, -2, 6, -2, -2, -2],
                           threshold=[1.5, 3.0, 0.3, -2, -2, 2.0, -2, -2, -2],
                           value=[0, 0, 0, 0, 1, 0, 3, 2, 1],
                           hsw_curve=(0.01, 0.02), max_acceptable_hsw_rate=0.5)


def _run_pair(H, O, T, cfg, lanes, batches, tree=None):
    """Feed the same per-chunk inputs [(pulses, counts, maxCalls, frames, lastFrame,
    blStats, score, pdStats, models)] to the GPU stage and to the oracle."""
    gm, om = H.Metrics(cfg, lanes), O.Metrics(cfg, lanes)
    if tree is not None:
        gm.set_activity_tree(tree)
        om.set_activity_tree(tree[0])
    d_out = H.dev_bytes(lanes * T.METRICS.itemsize)
    produced = 0
    for i, (p, c, mc, fr, last, bs, sc, pd, mo) in enumerate(batches):
        g = gm.process(H.dev_from_numpy(p), H.dev_from_numpy(c), mc, fr, last, H.dev_from_numpy(bs),
                       H.dev_from_numpy(sc), H.dev_from_numpy(pd), H.dev_from_numpy(mo), d_out)
        d, o = om.process(p, c, mc, fr, last, bs, sc, pd, mo, T.METRICS)
        assert g == bool(d), "chunk %d: block end" % i
        if d:
            produced += 1
            _eq_metrics(H.to_numpy(d_out, T.METRICS, (lanes,)), o, "chunk %d" % i)
    return produced


def test_metrics_populated_bit_exact(H, O, T, gpu):
    L = 3
    cfg = H.default_config()
    cfg.movie.frame_rate = MI.FRAME_RATE
    batches = [(*MI.pulses(L, b), 24, MI.FRAMES, (b + 1) * MI.FRAMES, MI.baseliner_stats(L), MI.viterbi_score(L),
                MI.pd_stats(L), MI.models(L)) for b in range(2 * MI.BATCHES)]
    assert _run_pair(H, O, T, cfg, L, batches) == 2
    assert _run_pair(H, O, T, cfg, L, batches, tree=_tree(H)) == 2


def _random_batches(T, L, frames, nchunks, seed, max_calls=32):
    rng = np.random.default_rng(seed)
    out = []
    t = np.zeros(L * 64, np.int64)
    for c in range(nchunks):
        p = np.zeros((L * 64, max_calls), dtype=T.PULSE)
        counts = rng.integers(0, max_calls + 3, size=L * 64).astype(np.uint32)   # some overflow the slots
        counts[: 64] = 0 if c % 3 == 0 else counts[: 64]                          # a lane with empty chunks
        for z in range(L * 64):
            n = min(int(counts[z]), max_calls)
            for j in range(n):
                t[z] += int(rng.integers(0, 3))
                w = int(rng.integers(1, 7))
                q = p[z, j]
                q["start"], q["width"] = t[z], w
                q["label"] = rng.integers(0, 4) if rng.random() > 0.05 else 5
                q["mid_signal"] = rng.integers(50, 4000)
                q["mean_signal"] = q["mid_signal"]
                q["max_signal"] = q["mid_signal"] + rng.integers(0, 400)
                q["signal_m2"] = np.float32(q["mid_signal"] * 0.1) ** 2 * max(w - 2, 0) * rng.uniform(1.0, 1.2)
                q["is_reject"] = rng.random() < 0.1
                t[z] += w
        bs = np.zeros(L, dtype=T.BASELINER_STATS)
        full = bs["full"]
        full["basic"]["m0"] = frames
        full["basic"]["m1"] = rng.normal(0, 50, (L, 64))
        full["basic"]["m2"] = rng.uniform(1e3, 1e5, (L, 64))
        full["lag_m2"] = rng.uniform(-1e3, 5e4, (L, 64))
        full["fbuf"] = rng.normal(0, 10, (L, 4, 64))
        full["bbuf"] = rng.normal(0, 10, (L, 4, 64))
        full["bidx"][:, 0, :] = 4
        full["bidx"][:, 1, :] = frames % 4
        bg = bs["background"]
        bg["offset"] = rng.normal(200, 5, (L, 64))
        bg["m0"] = rng.uniform(10, 400, (L, 64))
        bg["m1"] = rng.normal(0, 20, (L, 64))
        bg["m2"] = rng.uniform(100, 2000, (L, 64))
        pd = np.zeros(L, dtype=T.STAT_ACC)
        pd["offset"] = rng.normal(0, 3, (L, 64))
        pd["m0"] = rng.uniform(50, 500, (L, 64))
        pd["m0"][0, :8] = 0                                                        # ZMWs without baseline frames
        pd["m1"] = rng.normal(0, 30, (L, 64))
        pd["m2"] = rng.uniform(500, 5e4, (L, 64))
        mo = np.zeros(L, dtype=T.LANE_MODEL)
        mo["analogs"]["means"] = np.float16(rng.uniform(40, 400, (L, 4, 64))).view(np.uint16)
        mo["analogs"]["vars"] = np.float16(rng.uniform(5, 800, (L, 4, 64))).view(np.uint16)
        mo["baseline"]["vars"] = np.float16(30).view(np.uint16)
        score = rng.uniform(0, 3, L * 64).astype(np.float32)
        out.append((p.reshape(-1), counts, max_calls, frames, (c + 1) * frames, bs, score, pd, mo))
    return out


@pytest.mark.parametrize("seed", [3, 4])
def test_metrics_random_streams_bit_exact(H, O, T, gpu, seed):
    L, F = 4, 256
    cfg = H.default_config()
    cfg.frames_per_hf_metric_block = 1024
    batches = _random_batches(T, L, F, 9, seed)       # two blocks and a part of a third
    assert _run_pair(H, O, T, cfg, L, batches, tree=_tree(H)) == 2
    cfg.realtime_activity_labels = 0
    assert _run_pair(H, O, T, cfg, L, batches, tree=_tree(H)) == 2


def test_analyzer_metrics_end_to_end(H, O, T, gpu):
    """Analyzer with metrics blocks of 1024 frames over synthetic traces: the
    block flag and every metrics field against the CPU analyzer."""
    L, F, C = 4, 512, 9
    cfg = H.bench_config(F)
    cfg.frames_per_hf_metric_block = 1024
    chunks = gen_traces(O, T, H, L, F, C, seed=91)
    oa = O.Analyzer(cfg, 0, L, F, em="mirror")
    ga = H.BatchAnalyzer(0, L, F, cfg)
    tree = _tree(H)
    oa.set_activity_tree(tree[0])
    ga.set_activity_tree(tree)
    dts = (T.PULSE, T.STAT_ACC, T.BASELINER_STATS, T.LANE_MODEL)
    blocks = 0
    for c, tr in enumerate(chunks):
        r = oa(tr, c * F, dts, want=("metrics",))
        ga(tr, c * F)
        assert ga.flags() == r["flags"], "chunk %d flags" % c
        if r["flags"][3]:
            blocks += 1
            _eq_metrics(ga.metrics(), r["metrics"], "chunk %d" % c)
    assert blocks >= 2
