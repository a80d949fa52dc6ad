"""The GPU analyzer against the oracle's analyzer on the workloads the
pipeline tests do not reach at bench cadence: the stress distribution at
1024-frame chunks, the reference's default estimation cadence with its pool
stagger, and the production shape (32,768 lanes), where the last lanes of a
full-size run are compared with the oracle run on those lanes alone (lanes are
independent), so an offset that overflows in the large buffers would show."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DTS = None


def _dts(T):
    return (T.PULSE, T.STAT_ACC, T.BASELINER_STATS, T.LANE_MODEL)


def _compare(ga, r, what, c):
    assert ga.flags()[:3] == r["flags"][:3], (what, c)
    assert np.array_equal(ga.baselined(), r["baselined"]), (what, c)
    p, cnt, _ = ga.pulses()
    assert np.array_equal(cnt, r["counts"]), (what, c)
    if r["flags"][2]:
        assert ga.models().tobytes() == r["models"].tobytes(), (what, c)
        assert np.array_equal(ga.labels(), r["labels"]), (what, c)
    for z in np.nonzero(cnt)[0]:
        k = cnt[z]
        assert np.array_equal(np.ascontiguousarray(p[z, :k]).view(np.uint8),
                              np.ascontiguousarray(r["pulses"][z, :k]).view(np.uint8)), (what, c, z)


def test_stress_workload_f1024_vs_oracle(H, O, T, gpu):
    """The stress distribution (baseline sigma 25, amplitudes 90/60/40/25, a +60
    baseline step at a random frame, 5% empty ZMWs) at 1024-frame chunks: bit for
    bit against the mirror oracle every chunk, and labels against the reference
    form (every mismatch in a ZMW whose model differs)."""
    L, F, C = 8, 1024, 6
    gen = H.tracegen_default(seed=77, baseline_sigma=25.0, amplitude=(90.0, 60.0, 40.0, 25.0), drift_per_chunk=0.0,
                             step_drift=60.0, step_span=float(C * F), empty_fraction=0.05)
    st = np.zeros(L, dtype=T.TRACEGEN_STATE)
    cfg = H.bench_config(F)
    ga = H.BatchAnalyzer(0, L, F, cfg)
    om, orf = O.Analyzer(cfg, 0, L, F, em="mirror"), O.Analyzer(cfg, 0, L, F, em="reference")
    frames = mism = same = 0
    for c in range(C):
        tr = O.tracegen(gen, L, F, c * F, st, c == 0)
        ga(tr, c * F)
        r = om(tr, c * F, _dts(T), want=("pulses", "labels", "models", "baselined"))
        _compare(ga, r, "stress", c)
        rr = orf(tr, c * F, _dts(T), want=("labels", "models"))
        if rr["flags"][2]:
            zdiff = (r["models"].view(np.uint16).reshape(L, -1, 64) !=
                     rr["models"].view(np.uint16).reshape(L, -1, 64)).any(axis=1)
            d = r["labels"] != rr["labels"]
            frames += d.size
            mism += int(d.sum())
            same += int((d & ~zdiff[:, None, :]).sum())
    print("\nstress, 1024-frame chunks: GPU = mirror oracle bit for bit; labels vs reference form %.5f%% "
          "(%d mismatches, %d in ZMWs with equal models)" % (100 * (1 - mism / frames), mism, same))
    assert frames > 0 and same == 0


@pytest.mark.parametrize("pool", [1, 2, 3, 4, 5, 6, 7])
def test_reference_default_cadence_and_stagger_vs_oracle(H, O, T, gpu, pool):
    """The reference defaults (sigma EMA over 512 strides: 9216 frames of baseliner
    start-up; 1000 pre-accumulation and 4000 estimation frames, histograms
    accumulated over 8 chunks) with pool ids 1-7: the stagger (pool mod 8) x 512
    frames puts the first estimate at chunk 27 + pool.  Bit for bit against the
    mirror oracle every chunk, and the first estimate where the closed form says."""
    L, F, C = 2, 512, 44
    cfg = H.default_config()
    gen = H.tracegen_default(seed=900 + pool)
    st = np.zeros(L, dtype=T.TRACEGEN_STATE)
    ga = H.BatchAnalyzer(pool, L, F, cfg)
    oa = O.Analyzer(cfg, pool, L, F, em="mirror")
    est = []
    for c in range(C):
        tr = O.tracegen(gen, L, F, c * F, st, c == 0)
        ga(tr, c * F)
        r = oa(tr, c * F, _dts(T), want=("pulses", "labels", "models", "baselined"))
        _compare(ga, r, "pool %d" % pool, c)
        if r["flags"][1]:
            est.append(c)
    assert est[0] == 27 + pool, est
    assert all(b - a == 8 for a, b in zip(est, est[1:])), est


def _slice(dev_ptr, offset, nbytes, dtype, shape, H):
    out = np.empty(nbytes, dtype=np.uint8)
    H.check(H.lib.hwpa_memcpy_dtoh(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(dev_ptr + offset),
                                   ctypes.c_size_t(nbytes)))
    return out.view(dtype).reshape(shape)


def test_production_shape_last_lanes_vs_oracle(H, O, T, gpu):
    """32,768 lanes x 512 frames (2 GiB chunks, a 4.4 GB traceback): the last four
    lanes' baselined traces, labels, models and pulses equal the oracle run on
    those four lanes alone, through two estimates."""
    L, F, C, K = 32768, 512, 5, 4
    cfg = H.bench_config(F)
    gen = H.tracegen_default(seed=31)
    d_state = H.dev_bytes(L * T.TRACEGEN_STATE.itemsize)
    d_tr = H.dev_bytes(L * F * 64 * 2)
    ga = H.BatchAnalyzer(0, L, F, cfg)
    oa = O.Analyzer(cfg, 0, K, F, em="mirror")
    st = np.zeros(K, dtype=T.TRACEGEN_STATE)
    lane_bytes = F * 64 * 2
    off = (L - K) * lane_bytes
    n_est = 0
    for c in range(C):
        H.tracegen(gen, L, F, c * F, d_state, c == 0, d_tr)
        ga(d_tr, c * F)
        tr = O.tracegen(gen, K, F, c * F, st, c == 0, lane_offset=L - K)
        assert np.array_equal(_slice(d_tr.data_ptr(), off, K * lane_bytes, np.int16, (K, F, 64), H), tr)
        r = oa(tr, c * F, _dts(T), want=("pulses", "labels", "models", "baselined"))
        ga.sync()
        assert ga.flags()[:3] == r["flags"][:3], c
        p = ctypes.c_void_p()
        H.check(H.lib.hwpa_result_baselined(ga.h, ctypes.byref(p)))
        assert np.array_equal(_slice(p.value, off, K * lane_bytes, np.int16, (K, F, 64), H), r["baselined"]), c
        if r["flags"][2]:
            H.check(H.lib.hwpa_result_labels(ga.h, ctypes.byref(p)))
            assert np.array_equal(_slice(p.value, off, K * lane_bytes, np.int16, (K, F, 64), H), r["labels"]), c
            H.check(H.lib.hwpa_result_models(ga.h, ctypes.byref(p)))
            msz = T.LANE_MODEL.itemsize
            assert _slice(p.value, (L - K) * msz, K * msz, np.uint8, (K * msz,), H).tobytes() == \
                r["models"].tobytes(), c
        pp, pc, _, mc = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_uint32()
        H.check(H.lib.hwpa_result_pulses(ga.h, ctypes.byref(pp), ctypes.byref(pc), ctypes.byref(_),
                                         ctypes.byref(mc)))
        z0 = (L - K) * 64
        cnt = _slice(pc.value, z0 * 4, K * 64 * 4, np.uint32, (K * 64,), H)
        assert np.array_equal(cnt, r["counts"]), c
        ps = T.PULSE.itemsize
        pulses = _slice(pp.value, z0 * mc.value * ps, K * 64 * mc.value * ps, T.PULSE, (K * 64, mc.value), H)
        for z in np.nonzero(cnt)[0]:
            k = cnt[z]
            assert np.array_equal(np.ascontiguousarray(pulses[z, :k]).view(np.uint8),
                                  np.ascontiguousarray(r["pulses"][z, :k]).view(np.uint8)), (c, z)
        n_est += r["flags"][1]
    assert n_est >= 2
