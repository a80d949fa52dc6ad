"""The DME oracle in its two forms (CPU).

"reference" follows DmeEmHost::EstimateLaneDetModel as written
(traceAnalysis/DmeEmHost.cpp:216-615): log-sum-exp E-step with exp and a
division per bin boundary, the reference's Cephes exp/log and libm
exp2/log2/pow.  "mirror" is the GPU kernel's restated E-step (densities carried
directly, per-boundary fused sums, density recurrence, table log), which the GPU
matches bit for bit.  The two are algebraically different formulations of the
same EM, so their agreement pins both; where they part, the cause must be the
one the reference's capture rule amplifies: each ZMW's estimate is the
parameter set of the LAST iteration whose log-likelihood change fell in
[0, 1e-4 |LL|) (DmeEmHost.cpp:470-483), and a change near 0 flips with the last
bit of LL.  "reference_libm" (the reference form with libm expf/logf in place
of the Cephes routines) measures that floor: two faithful restatements that
differ only in the last bits of exp and log.
"""
import numpy as np
import pytest

from test_gpu_parity import _em_case

LANES = 16
FI = ((512, 1024), (1024, 1536))


def _run(O, cfg, h, s, m, form):
    models, diag = O.dme_estimate(cfg, h, s, m, FI[0], FI[1], want_diag=True, form=form)
    L = len(models)
    est = np.array([np.array(diag[l].est_fp32) for l in range(L)])      # [lane][mode][w, mean, var][zmw]
    out = np.array([np.array(diag[l].out_fp32) for l in range(L)])
    cap = np.array([np.array(diag[l].capture_iter) for l in range(L)])   # [lane][zmw]
    its = np.array([diag[l].iterations_run for l in range(L)])
    return models, est, out, cap, its


def _zmw_rel(a, b):
    """Largest relative difference over the 15 parameters of each ZMW: [lane][zmw].
    Means are taken relative to max(|mean|, baseline sigma): the baseline mean of
    baselined traces sits at 0 +- noise, where a plain relative error has no scale
    (values of 4e-5 counts differing by 1e-6)."""
    scale = np.ones_like(b)
    scale[:, :, 1, :] = np.sqrt(np.abs(b[:, 0:1, 2, :]))
    den = np.maximum(np.abs(b), np.where(np.arange(3)[None, None, :, None] == 1, scale, 1e-30))
    return (np.abs(a - b) / den).max(axis=(1, 2))


@pytest.fixture(scope="module")
def pipeline_case(H, O, T):
    return _em_case(H, O, T, LANES, 512, seed=11)


def test_forms_agree_within_1e4_except_capture_flips(O, pipeline_case):
    cfg, h, s, m = pipeline_case
    _, e_ref, o_ref, c_ref, it_ref = _run(O, cfg, h, s, m, "reference")
    _, e_mir, o_mir, c_mir, it_mir = _run(O, cfg, h, s, m, "mirror")
    assert np.array_equal(it_ref, it_mir)                 # the lane-wide exit agrees
    rel = _zmw_rel(e_mir, e_ref)
    flip = c_ref != c_mir
    # ZMWs whose estimate was captured at the same iteration agree within 1e-4 on every parameter
    assert rel[~flip].max() <= 1e-4, rel[~flip].max()
    # every ZMW outside 1e-4 is a capture-iteration flip, and those are few
    assert not np.any((rel > 1e-4) & ~flip)
    assert flip.mean() < 0.05, flip.mean()
    assert (rel <= 1e-4).mean() > 0.95


def test_libm_yardstick_flips_like_the_mirror(O, pipeline_case):
    """The reference form against itself with libm exp/log: the same kind of
    capture flips at a similar rate, so no implementation without the
    reference's exact bits can hold 1e-4 on every ZMW."""
    cfg, h, s, m = pipeline_case
    _, e_ref, _, c_ref, _ = _run(O, cfg, h, s, m, "reference")
    _, e_lib, _, c_lib, _ = _run(O, cfg, h, s, m, "reference_libm")
    rel = _zmw_rel(e_lib, e_ref)
    flip = c_ref != c_lib
    assert flip.any()
    assert not np.any((rel > 1e-4) & ~flip)
    assert rel[~flip].max() <= 1e-4


@pytest.mark.parametrize("form", ["reference", "mirror"])
def test_em_expected_counts_recover_model(H, O, T, form):
    """A histogram holding the expected bin counts of a mixture that satisfies
    the EM's own model (analog means s * relAmp, analog variances
    blVar + mean + cv^2 mean^2, ModelSignalCovar with an empty PSF) is
    recovered to within the binning error: the grouped-data EM's fixed point
    (McLachlan & Jones 1988, cited at DmeEmHost.cpp:346-349), reached here by
    iterating to a high limit (EmIterationLimit, IterateToLimit)."""
    from math import erf, sqrt
    cfg = H.default_config()
    sigma, s_true = 10.0, 20.0
    rel = np.array([3.6, 2.4, 1.6, 1.0])
    means = np.concatenate([[1.5], s_true * rel])
    var = np.concatenate([[sigma ** 2], sigma ** 2 + means[1:] + 0.01 * means[1:] ** 2])
    w = np.array([0.6, 0.1, 0.1, 0.1, 0.1])
    cfg.movie.ref_snr = s_true / sigma
    # iterate to a high limit so the last captured iteration sits at the fixed point
    cfg.em_iteration_limit = 300
    cfg.iterate_to_limit = 1
    n_total = 60000
    lo, bs = -60.0, 1.0
    edges = lo + bs * np.arange(301)

    def cdf(x):
        return sum(w[k] * 0.5 * (1.0 + erf((x - means[k]) / sqrt(2.0 * var[k]))) for k in range(5))

    c = np.array([cdf(x) for x in edges])
    counts = np.round(np.diff(c) * n_total).astype(np.int64)
    hist = np.zeros(1, dtype=T.LANE_HIST)
    stats = np.zeros(1, dtype=T.BASELINER_STATS)
    hist["low_bound"][0, :], hist["bin_size"][0, :] = lo, bs
    hist["bin_count"][0, :, :] = counts[:, None]
    hist["outlier_low"][0, :] = int(round(c[0] * n_total))
    hist["outlier_high"][0, :] = int(round((1.0 - c[-1]) * n_total))
    nb = n_total * w[0]
    stats["baseline"]["m0"][0, :] = nb
    stats["baseline"]["m1"][0, :] = nb * means[0]
    stats["baseline"]["m2"][0, :] = nb * (var[0] + means[0] ** 2)
    stats["full"]["basic"]["m0"][0, :] = n_total
    models = O.dme_init_models(cfg, stats, T.LANE_MODEL)
    _, diag = O.dme_estimate(cfg, hist, stats, models, (0, 4000), (0, 4000), want_diag=True, form=form)
    est = np.array(diag[0].est_fp32)[:, :, 0]         # [mode][w, mean, var]
    assert diag[0].converged[0]
    np.testing.assert_allclose(est[:, 0], w, atol=2e-3)
    np.testing.assert_allclose(est[1:, 1], means[1:], rtol=2e-3)
    assert abs(est[0, 1] - means[0]) < 0.05 * sigma
    # the baseline variance carries the quantisation term binSize^2 / 12
    np.testing.assert_allclose(est[0, 2], var[0] + bs * bs / 12.0, rtol=2e-2)
    np.testing.assert_allclose(est[1:, 1] / rel, est[1, 1] / rel[0], rtol=1e-6)


def test_forms_over_several_seeds(H, O, T):
    """Five seeded cases of 48 lanes (15,360 ZMWs), the mirror form (the GPU's bits) and
    the libm yardstick each against the reference form.  Beyond capture flips, a few
    ZMWs whose capture iteration matches still differ by a little over 1e-4: slowly
    converging ZMWs carry the last-bit differences of the E-step through 13-19
    iterations.  The yardstick, two faithful restatements apart, shows the same kind."""
    nb = {"mirror": 0, "reference_libm": 0}
    mx = {"mirror": 0.0, "reference_libm": 0.0}
    within = flips = total = 0
    for seed in (5, 9, 11, 21, 33):
        cfg, h, s, m = _em_case(H, O, T, 48, 512, seed=seed)
        _, e_ref, _, c_ref, it_ref = _run(O, cfg, h, s, m, "reference")
        total += c_ref.size
        for form in nb:
            _, e, _, c, it = _run(O, cfg, h, s, m, form)
            assert np.array_equal(it, it_ref)
            rel = _zmw_rel(e, e_ref)
            flip = c != c_ref
            nb[form] += int(((rel > 1e-4) & ~flip).sum())
            mx[form] = max(mx[form], float(rel[~flip].max()))
            if form == "mirror":
                within += int((rel <= 1e-4).sum())
                flips += int(flip.sum())
    print("\nmirror: %.3f%% within 1e-4, %.2f%% flips, %d same-capture above 1e-4 (max %.2e); yardstick %d (max %.2e)"
          % (100.0 * within / total, 100.0 * flips / total, nb["mirror"], mx["mirror"], nb["reference_libm"],
             mx["reference_libm"]))
    # measured: 97.82% within, 2.32% flips, 3 same-capture ZMWs above 1e-4 (max 3.2e-4); yardstick 1 (1.5e-4)
    assert within / total > 0.97
    assert flips / total < 0.03
    assert nb["mirror"] <= 0.0005 * total and mx["mirror"] < 5e-4
    assert nb["reference_libm"] >= 1      # the floor exists between faithful restatements too
