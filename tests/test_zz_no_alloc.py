"""The learned-start estimation without allocation
(hwpa_dme_estimate_learned_noalloc) against the allocating entry point
(hwpa_dme_estimate_learned) on the same inputs, bit for bit: models, EM
iteration counts and fallback counts.  Two calls in a row, the second on
fewer lanes, so a workspace that kept the first call's contents would show.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

F = 512


@pytest.fixture(scope="module")
def net(H, gpu):
    from hw_amd import learned as Lr
    pairs = Lr.generate_pairs(64, 6, seed=211, snr=(4.0, 20.0), min_conf=0.0)
    init = H.Initializer(seed=3)
    Lr.train(init, pairs, epochs=6, batch=1024, lr=2e-3)
    return init


def _case(H, O, T, L, seed):
    from test_gpu_parity import _em_case
    cfg, h, s, m = _em_case(H, O, T, L, F, seed=seed, tg=dict(snr_lo=4.0, snr_hi=20.0))
    cfg.learned_init = 1
    return cfg, h, s, m


def _run(H, T, fn, cfg, net, h, s, m, L):
    mfi, efi = (512, 1024), (1024, 1536)
    d_models = H.dev_from_numpy(m)
    d_iters = H.dev_bytes(L * 64 * 4)
    n = fn(cfg, net, H.dev_from_numpy(h), H.dev_from_numpy(s), d_models, mfi, efi, L, d_iters)
    return n, H.to_numpy(d_models, T.LANE_MODEL, (L,)).tobytes(), H.to_numpy(d_iters, np.int32, (L,))


def test_noalloc_equals_allocating_entry_point(H, O, T, net):
    for L, seed in ((40, 901), (9, 902)):
        print("noalloc case: %d lanes, seed %d" % (L, seed), flush=True)
        cfg, h, s, m = _case(H, O, T, L, seed)
        n_ref, mod_ref, it_ref = _run(H, T, H.dme_estimate_learned, cfg, net, h, s, m, L)
        n_na, mod_na, it_na = _run(H, T, H.dme_estimate_learned_noalloc, cfg, net, h, s, m, L)
        print("  fallback %d / %d ZMWs" % (n_ref, L * 64))
        assert n_na == n_ref
        assert mod_na == mod_ref
        assert np.array_equal(it_na, it_ref)
        assert 0 < n_ref < L * 64          # accepted and fallback paths both ran


def test_noalloc_rejects_more_lanes_than_its_workspace(H, T, net, gpu):
    L = H.NOALLOC_MAX_LANES + 1
    cfg = H.bench_config(F)
    cfg.learned_init = 1
    d = H.dev_bytes(256)      # never read: the lane count is rejected before any launch
    with pytest.raises(Exception, match="fixed workspace"):
        H.dme_estimate_learned_noalloc(cfg, net, d, d, d, (0, 512), (512, 1024), L)
