"""The GPU path against the reference-form oracle (DmeEmHost's own arithmetic).

The GPU's EM restates the E-step (test_dme_forms.py describes how) and equals
the "mirror" oracle bit for bit; here it is measured against the "reference"
oracle at the tolerances the path is held to: detection-model parameters within
1e-4 relative, frame labels on at least 99.99% of frames.  Every ZMW outside
1e-4 must be a capture-iteration flip (the estimate taken at a different
iteration, see test_dme_forms.py) and every label mismatch must lie in a ZMW
whose half-precision model differs; with equal models the labels, baselined
traces and pulses are identical.
"""
import numpy as np
import pytest

from conftest import gen_traces
from test_dme_forms import FI, _run, _zmw_rel
from test_gpu_parity import _dev, _em_case

pytestmark = pytest.mark.gpu


def _gpu_est(H, T, cfg, h, s, m):
    L = len(m)
    d_models = _dev(H, m)
    d_diag = H.dev_bytes(L * H.EM_DIAG_ROWS * T.LANE * 4)
    H.dme_estimate_diag(cfg, _dev(H, h), _dev(H, s), d_models, FI[0], FI[1], L, d_diag)
    diag = H.to_numpy(d_diag, np.float32, (L, H.EM_DIAG_ROWS, T.LANE))
    # rows 0-4 weights, 5-9 means, 10-14 variances -> [lane][mode][w, mean, var][zmw]
    est = diag[:, :15].reshape(L, 3, 5, T.LANE).transpose(0, 2, 1, 3)
    return H.to_numpy(d_models, T.LANE_MODEL, (L,)), est


def test_gpu_em_within_1e4_of_reference_form(H, O, T, gpu):
    """64 lanes x 3 estimates (12,288 ZMWs) of the bench workload."""
    L = 64
    rels, flips = [], []
    for seed in (11, 12, 13):
        cfg, h, s, m = _em_case(H, O, T, L, 512, seed=seed)
        g_models, g_est = _gpu_est(H, T, cfg, h, s, m)
        m_models, m_est, _, m_cap, _ = _run(O, cfg, h, s, m, "mirror")
        assert np.array_equal(g_est, m_est), "GPU EM differs from the mirror oracle (seed %d)" % seed
        assert np.array_equal(g_models.view(np.uint8), m_models.view(np.uint8))
        _, r_est, _, r_cap, _ = _run(O, cfg, h, s, m, "reference")
        rels.append(_zmw_rel(g_est, r_est))
        flips.append(m_cap != r_cap)
    rel, flip = np.concatenate(rels), np.concatenate(flips)
    print("\nGPU EM vs reference form: %d ZMWs, %.4f%% within 1e-4 on all 15 parameters, %.3f%% capture flips, "
          "largest same-capture difference %.2e, quantiles 50/99/99.9%% %s"
          % (rel.size, 100 * (rel <= 1e-4).mean(), 100 * flip.mean(), rel[~flip].max(),
             np.array2string(np.quantile(rel, [0.5, 0.99, 0.999]), precision=2)))
    assert rel[~flip].max() <= 1e-4
    assert not np.any((rel > 1e-4) & ~flip)
    assert flip.mean() < 0.05
    assert (rel <= 1e-4).mean() > 0.95


def test_gpu_analyzer_labels_vs_reference_form(H, O, T, gpu):
    """64 lanes x 16 chunks end to end (bench configuration, an estimate every
    chunk): labels agree on >= 99.99% of frames, all mismatches in ZMWs whose
    model differs; baselined traces bit-exact every chunk; pulses bit-exact in
    every ZMW whose models and labels have been equal in every chunk so far."""
    L, F, C = 64, 512, 16
    cfg = H.bench_config(F)
    chunks = gen_traces(O, T, H, L, F, C, seed=21)
    oa = O.Analyzer(cfg, 0, L, F, em="reference")
    ga = H.BatchAnalyzer(0, L, F, cfg)
    dts = (T.PULSE, T.STAT_ACC, T.BASELINER_STATS, T.LANE_MODEL)
    frames = mism = mism_same_model = model_diff_zmws = 0
    clean = np.ones(L * T.LANE, dtype=bool)     # ZMWs whose models and labels have been equal so far
    for c, tr in enumerate(chunks):
        r = oa(tr, c * F, dts, want=("pulses", "labels", "models", "baselined"))
        ga(tr, c * F)
        assert ga.flags()[:3] == r["flags"][:3]
        assert np.array_equal(ga.baselined(), r["baselined"]), "chunk %d baselined" % c
        if not r["flags"][2]:
            continue
        gm = ga.models().view(np.uint16).reshape(L, -1, T.LANE)
        om = r["models"].view(np.uint16).reshape(L, -1, T.LANE)
        zdiff = (gm != om).any(axis=1)
        lab = ga.labels()
        d = lab != r["labels"]
        frames += d.size
        mism += int(d.sum())
        mism_same_model += int((d & ~zdiff[:, None, :]).sum())
        model_diff_zmws += int(zdiff.sum())
        p, cnt, _ = ga.pulses()
        clean &= (~zdiff & ~d.any(axis=1)).reshape(-1)
        same = clean
        assert np.array_equal(cnt[same], r["counts"][same]), "chunk %d pulse counts" % c
        for z in np.nonzero(same & (cnt > 0))[0]:
            k = cnt[z]
            assert np.array_equal(np.ascontiguousarray(p[z, :k]).view(np.uint8),
                                  np.ascontiguousarray(r["pulses"][z, :k]).view(np.uint8)), "chunk %d zmw %d" % (c, z)
    agree = 1.0 - mism / frames
    print("\nGPU analyzer vs reference form: %d labelled frames, agreement %.6f%%, %d mismatches "
          "(%d in ZMWs with equal models), %d ZMW-chunks with a differing model"
          % (frames, 100 * agree, mism, mism_same_model, model_diff_zmws))
    assert mism_same_model == 0
    assert agree >= 0.9999
