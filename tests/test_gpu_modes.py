"""GPU parity of the analyzer's model-estimation modes (Fixed, Initial,
Dynamic) and of uint8 traces against the CPU oracle, through the C-ABI,
plus a batch whose pools mix modes and encodings against stand-alone pools."""
import numpy as np
import pytest

from conftest import gen_traces
from test_gpu_metrics import _eq_metrics

pytestmark = pytest.mark.gpu


def _dts(T):
    return (T.PULSE, T.STAT_ACC, T.BASELINER_STATS, T.LANE_MODEL)


def _u8(tr):
    return np.clip(tr.astype(np.int32) // 2 + 64, 0, 255).astype(np.uint8)


def test_baseliner_uint8_bit_exact(H, O, T, gpu):
    cfg = H.default_config()
    cfg.movie.pedestal = 64
    L, F = 4, 512
    b8, b16 = H.Baseliner(cfg, L), H.Baseliner(cfg, L)
    d_o8, d_o16 = H.dev_bytes(L * F * 64 * 2), H.dev_bytes(L * F * 64 * 2)
    d_s8, d_s16 = H.dev_bytes(L * T.BASELINER_STATS.itemsize), H.dev_bytes(L * T.BASELINER_STATS.itemsize)
    for c, tr in enumerate(gen_traces(O, T, H, L, F, 3, seed=31)):
        u = _u8(tr)
        b8.run(H.dev_from_numpy(u), F, d_o8, d_s8, encoding=T.ENCODING_UINT8)
        b16.run(H.dev_from_numpy(u.astype(np.int16)), F, d_o16, d_s16)
        assert np.array_equal(H.to_numpy(d_o8, np.int16, (L, F, 64)), H.to_numpy(d_o16, np.int16, (L, F, 64))), c
        assert H.to_numpy(d_s8, np.uint8, (L * T.BASELINER_STATS.itemsize,)).tobytes() == \
            H.to_numpy(d_s16, np.uint8, (L * T.BASELINER_STATS.itemsize,)).tobytes(), c


@pytest.mark.parametrize("mode,encoding", [(0, 0), (0, 1), (1, 0), (1, 1), (2, 1)])
def test_analyzer_modes_end_to_end(H, O, T, gpu, mode, encoding):
    L, F, C = 4, 512, 8
    cfg = H.bench_config(F)
    cfg.model_estimation_mode = mode
    cfg.frames_per_hf_metric_block = 1024
    chunks = gen_traces(O, T, H, L, F, C, seed=61 + mode)
    if encoding:
        cfg.movie.pedestal = 64
        chunks = [_u8(tr) for tr in chunks]
    oa = O.Analyzer(cfg, 0, L, F, em="mirror")
    ga = H.BatchAnalyzer(0, L, F, cfg)
    seen = [0, 0, 0, 0]
    for c, tr in enumerate(chunks):
        r = oa(tr, c * F, _dts(T), want=("pulses", "labels", "models", "baselined", "metrics"))
        ga(tr, c * F)
        f = ga.flags()
        assert f == r["flags"], "chunk %d flags %s vs %s" % (c, f, r["flags"])
        seen = [s + x for s, x in zip(seen, f)]
        assert np.array_equal(ga.baselined(), r["baselined"]), c
        if f[2]:
            assert ga.models().tobytes() == r["models"].tobytes(), c
            assert np.array_equal(ga.labels(), r["labels"]), c
        p, n, _ = ga.pulses()
        assert np.array_equal(n, r["counts"]), c
        for z in np.nonzero(n)[0]:
            assert p[z, :n[z]].tobytes() == r["pulses"][z, :n[z]].tobytes(), (c, z)
        if f[3]:
            _eq_metrics(ga.metrics(), r["metrics"], "chunk %d" % c)
    assert seen[0] >= 3 and seen[3] >= 2


def test_batch_mixed_modes_and_encodings(H, O, T, gpu):
    """One graph per step over pools of every mode, int16 and uint8, equals the
    stand-alone analyzers."""
    specs = []
    for i, (mode, enc) in enumerate([(0, 1), (1, 0), (2, 1), (2, 0)]):
        cfg = H.bench_config(512)
        cfg.model_estimation_mode = mode
        cfg.frames_per_hf_metric_block = 1024
        if enc:
            cfg.movie.pedestal = 64
        specs.append((20 + i, 3, 512, cfg, enc))
    steps = 7
    chunks = []
    for i, (_, L, F, _, enc) in enumerate(specs):
        host = gen_traces(O, T, H, L, F, steps, seed=70 + i)
        chunks.append([H.dev_from_numpy(_u8(t) if enc else t) for t in host])
    alone = []
    for i, (pid, L, F, cfg, enc) in enumerate(specs):
        a = H.BatchAnalyzer(pid, L, F, cfg)
        out = []
        for c in range(steps):
            a(chunks[i][c], c * F, encoding=enc)
            out.append((a.flags(), a.baselined().copy(), a.pulses()[1].copy(), a.models().tobytes()))
        alone.append(out)
    b = H.Batch([s[:4] for s in specs])
    for c in range(steps):
        b.process([chunks[i][c] for i in range(len(specs))], [c * 512] * len(specs), encodings=[s[4] for s in specs])
        b.sync()
        for i in range(len(specs)):
            a = b.pool(i)
            fl, bl, cnt, mo = alone[i][c]
            assert a.flags() == fl, (i, c)
            assert np.array_equal(a.baselined(), bl), (i, c)
            assert np.array_equal(a.pulses()[1], cnt), (i, c)
            if fl[2]:
                assert a.models().tobytes() == mo, (i, c)
