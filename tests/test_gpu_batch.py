"""Batched pools on the GPU: one captured hipGraph per chunk step must give
every pool exactly the results of its own stand-alone analyzer (same
inputs), with and without graphs, including pools whose stage sequences
differ from step to step (reference-default estimation interval and the
pool stagger), and the learned-start estimator inside the graph."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STEPS = 9


def _pools(H):
    cfg_def = H.default_config()
    cfg_def.sigma_ema_scale_strides = 0.0          # short start-up, estimates every 4000 frames
    pools = [(7, 5, 256, H.bench_config(256)), (8, 3, 512, H.bench_config(512)),
             (9, 4, 1024, H.bench_config(1024)), (10, 2, 512, cfg_def)]
    for p in pools:
        p[3].frames_per_hf_metric_block = 1024    # metrics blocks end inside the run (graph replays with resets)
    return pools


def _chunks(H, T, pools):
    out = []
    for i, (_, L, F, _) in enumerate(pools):
        gen = H.tracegen_default(seed=500 + i)
        st = H.dev_bytes(L * T.TRACEGEN_STATE.itemsize)
        bufs = []
        for c in range(STEPS):
            b = H.dev_bytes(L * F * 64 * 2)
            H.tracegen(gen, L, F, c * F, st, c == 0, b)
            bufs.append(b)
        out.append(bufs)
    H.sync()
    return out


def _snapshot(a):
    counts = a.pulses()[1]
    f = a.flags()
    return {"flags": f, "labels": a.labels(), "models": a.models().tobytes(),
            "baselined": a.baselined(), "counts": counts, "metrics": a.metrics().tobytes() if f[3] else None}


def _alone(H, pools, chunks, init=None):
    res = []
    for i, (pid, L, F, cfg) in enumerate(pools):
        an = H.BatchAnalyzer(pid, L, F, cfg)
        if init is not None:
            an.set_initializer(init)
        res.append([_snapshot(an(chunks[i][c], c * F)) for c in range(STEPS)])
    return res


def _batched(H, pools, chunks, graph, init=None):
    b = H.Batch(pools)
    b.set_graph(graph)
    if init is not None:
        b.set_initializer(init)
    res = [[] for _ in pools]
    for c in range(STEPS):
        b.process([chunks[i][c] for i in range(len(pools))], [c * p[2] for p in pools])
        b.sync()
        for i in range(len(pools)):
            res[i].append(_snapshot(b.pool(i)))
    return res, b.graph_stats()


def _same(x, y):
    for i in range(len(x)):
        for c in range(STEPS):
            a, b = x[i][c], y[i][c]
            assert a["flags"] == b["flags"], (i, c)
            assert np.array_equal(a["baselined"], b["baselined"]), (i, c)
            assert np.array_equal(a["counts"], b["counts"]), (i, c)
            if a["flags"][2]:      # labels (and so models) exist after the start-up window
                assert a["models"] == b["models"], (i, c)
                assert np.array_equal(a["labels"], b["labels"]), (i, c)
            assert a["metrics"] == b["metrics"], (i, c)
    assert sum(x[i][c]["flags"][2] for i in range(len(x)) for c in range(STEPS)) >= len(x) * 3
    assert sum(x[i][c]["flags"][3] for i in range(len(x)) for c in range(STEPS)) >= 4


@pytest.mark.parametrize("graph", [True, False])
def test_batch_matches_standalone_pools(H, T, gpu, graph):
    pools = _pools(H)
    chunks = _chunks(H, T, pools)
    ref = _alone(H, pools, chunks)
    got, stats = _batched(H, pools, chunks, graph)
    _same(ref, got)
    assert stats["steps"] == STEPS
    if graph:
        # the pools' stage sequences change only during start-up and at estimates
        assert stats["instantiations"] < STEPS
        assert stats["updates"] >= 1


def test_batch_learned_start_in_graph(H, T, gpu):
    pools = _pools(H)[:3]
    for p in pools:
        p[3].learned_init = 1
    chunks = _chunks(H, T, pools)
    init = H.Initializer(seed=3)
    ref = _alone(H, pools, chunks, init=init)
    got, _ = _batched(H, pools, chunks, True, init=init)
    _same(ref, got)


def test_batch_pools_match_oracle(H, O, T, gpu):
    """Every pool of the captured batch against the oracle's analyzer of that pool,
    bit for bit (mirror EM), every step: 256-, 512- and 1024-frame pools at the
    bench cadence, and a pool at the reference's estimation interval (1000
    pre-accumulation + 4000 estimation frames) with pool id 10, whose stagger is
    (10 mod 8) x 512 frames, run long enough for its first estimates."""
    steps = 18
    pools = _pools(H)
    res = [[] for _ in pools]
    b = H.Batch(pools)
    b.set_graph(True)
    gens = [(H.tracegen_default(seed=700 + i), H.dev_bytes(L * T.TRACEGEN_STATE.itemsize))
            for i, (_, L, _, _) in enumerate(pools)]
    bufs = [H.dev_bytes(L * F * 64 * 2) for (_, L, F, _) in pools]
    oas = [O.Analyzer(cfg, pid, L, F, em="mirror") for (pid, L, F, cfg) in pools]
    dts = (T.PULSE, T.STAT_ACC, T.BASELINER_STATS, T.LANE_MODEL)
    estimates = [0] * len(pools)
    for c in range(steps):
        for i, (_, L, F, _) in enumerate(pools):
            H.tracegen(gens[i][0], L, F, c * F, gens[i][1], c == 0, bufs[i])
        H.sync()
        b.process(bufs, [c * p[2] for p in pools])
        b.sync()
        for i, (pid, L, F, cfg) in enumerate(pools):
            tr = H.to_numpy(bufs[i], np.int16, (L, F, 64))
            r = oas[i](tr, c * F, dts, want=("pulses", "labels", "models", "baselined"))
            g = _snapshot(b.pool(i))
            assert g["flags"][:3] == r["flags"][:3], (pid, c)
            assert np.array_equal(g["baselined"], r["baselined"]), (pid, c)
            assert np.array_equal(g["counts"], r["counts"]), (pid, c)
            if r["flags"][2]:
                assert g["models"] == r["models"].tobytes(), (pid, c)
                assert np.array_equal(g["labels"], r["labels"]), (pid, c)
            p = b.pool(i).pulses()[0]
            for z in np.nonzero(g["counts"])[0]:
                k = g["counts"][z]
                assert np.array_equal(np.ascontiguousarray(p[z, :k]).view(np.uint8),
                                      np.ascontiguousarray(r["pulses"][z, :k]).view(np.uint8)), (pid, c, z)
            estimates[i] += r["flags"][1]
    # the bench-cadence pools estimate every chunk after start-up; the reference-cadence
    # pool (start-up 2 chunks + stagger 2 + 2 pre-accumulation + 8 estimation) from step 13
    assert all(n >= 10 for n in estimates[:3]) and estimates[3] >= 1, estimates
