"""Writes tests/golden/analyzer_digests.json: SHA-256 digests of the CPU
oracle's outputs (traces, baselined traces, labels, pulses, models) for a
fixed seeded case, chunk by chunk.  Rerun only when the oracle is meant to
change; tests/test_golden.py checks the oracle still reproduces them."""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

CASE = {"lanes": 2, "frames": 512, "chunks": 6, "seed": 20240611, "pool_id": 0}


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).view(np.uint8).tobytes()).hexdigest()


def run_case():
    import hw_amd as H
    import oracle as O
    from hw_amd import types as T
    L, F, C = CASE["lanes"], CASE["frames"], CASE["chunks"]
    cfg = H.bench_config(F)
    p = H.tracegen_default(seed=CASE["seed"])
    st = np.zeros(L, dtype=T.TRACEGEN_STATE)
    an = O.Analyzer(cfg, CASE["pool_id"], L, F, em="mirror")
    dts = (T.PULSE, T.STAT_ACC, T.BASELINER_STATS, T.LANE_MODEL)
    out = []
    for c in range(C):
        tr = O.tracegen(p, L, F, c * F, st, c == 0)
        r = an(tr, c * F, dts, want=("pulses", "labels", "models", "baselined", "bl_stats"))
        valid = np.concatenate([r["pulses"][z, :r["counts"][z]] for z in range(L * 64)])
        out.append({"chunk": c, "flags": r["flags"][:3], "traces": digest(tr), "baselined": digest(r["baselined"]),
                    "bl_stats": digest(r["bl_stats"]), "labels": digest(r["labels"]), "models": digest(r["models"]),
                    "pulses": digest(valid), "pulse_count": int(r["counts"].sum())})
    return out


if __name__ == "__main__":
    with open(os.path.join(HERE, "analyzer_digests.json"), "w") as f:
        json.dump({"case": CASE, "chunks": run_case()}, f, indent=1)
    print("wrote analyzer_digests.json")
