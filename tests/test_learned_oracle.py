"""The oracle's learned-start bookkeeping, on the CPU.

The learned start has no counterpart in the reference; what it must keep from the
reference is the plain EM (DmeEmHost::EstimateLaneDetModel, DmeEmHost.cpp:216-615)
for every ZMW whose capped EM is rejected.  That EM treats each ZMW on its own
(converged ZMWs are frozen while their lane iterates on), so a rejected ZMW must
end with exactly the model the plain oracle gives it in its original lane, whether
its lane fell back whole (plain EM in place) or in part (packed into compact lanes
and scattered back).  A wrong compaction offset, a scatter to the wrong slot, a
lost pad or a whole lane left out would each break this for some ZMW."""
import numpy as np

from test_gpu_parity import _em_case

FIELDS = ("means", "vars", "weights")


def _zmw(model, z):
    v = [model["confidence"][z]] + [model["baseline"][f][z] for f in FIELDS]
    v += [model["analogs"][f][k][z] for k in range(4) for f in FIELDS]
    return np.array(v, dtype=np.uint16)


def test_rejected_zmws_end_with_the_plain_estimate(H, O, T):
    L, F = 12, 512
    cfg, h, s, m = _em_case(H, O, T, L, F, seed=404, tg=dict(snr_lo=4.0, snr_hi=20.0))
    # lane 5: 100 frames per ZMW (insufficient data), so every ZMW is rejected after the capped EM
    h["bin_count"][5] = 0
    h["bin_count"][5, 10] = 100
    h["outlier_low"][5] = 0
    h["outlier_high"][5] = 0
    cfg.learned_init = 1
    mfi, efi = (512, 1024), (1024, 1536)
    plain, _ = O.dme_estimate(cfg, h, s, m, mfi, efi, want_diag=True, form="mirror")

    # Network outputs that start each ZMW at its plain estimate (identity output map), so most
    # converge within the cap and are accepted; lane 2 is below the confidence gate whole.
    f16 = lambda a: np.asarray(a).view(np.float16).astype(np.float64)
    lb = h["low_bound"].astype(np.float64)
    u = 2.0 * h["bin_size"].astype(np.float64)
    with np.errstate(all="ignore"):
        y = np.zeros((L, 64, 12))
        y[:, :, 0] = (f16(plain["baseline"]["means"]) - lb) / u
        y[:, :, 1] = np.log(np.maximum(f16(plain["baseline"]["vars"]), 1e-3) / u ** 2)
        y[:, :, 6] = np.log(np.maximum(f16(plain["baseline"]["weights"]), 1e-3))
        for a in range(4):
            y[:, :, 2 + a] = (f16(plain["analogs"]["means"][:, a]) - lb) / u
            y[:, :, 7 + a] = np.log(np.maximum(f16(plain["analogs"]["weights"][:, a]), 1e-3))
    y[:, :, 11] = 10.0
    y[2, :, 11] = -10.0
    y = np.nan_to_num(y, nan=0.0, posinf=0.0, neginf=0.0).reshape(L * 64, 12).astype(np.float32)
    omap = np.concatenate([np.zeros(6, np.float32), np.ones(6, np.float32)])

    learned, fb, n = O.dme_estimate_learned(cfg, h, s, m, mfi, efi, y, omap)
    per_lane = fb.reshape(L, 64).astype(np.int32).sum(axis=1)
    assert n == int(per_lane.sum())
    assert per_lane[2] == 64 and per_lane[5] == 64           # whole lanes: gated, and rejected
    assert ((per_lane > 0) & (per_lane < 64)).sum() >= 5     # lanes compacted in part
    assert n < L * 64 - 64                                   # and accepted ZMWs
    for l in range(L):
        for z in range(64):
            if fb[l * 64 + z]:
                assert np.array_equal(_zmw(learned[l], z), _zmw(plain[l], z)), (l, z)
