"""Regression pin: the CPU oracle reproduces the committed digests of
tests/golden/analyzer_digests.json (written by tests/golden/make_golden.py)."""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))


def test_oracle_reproduces_golden_digests():
    import sys
    sys.path.insert(0, os.path.join(HERE, "golden"))
    import make_golden
    ref = json.load(open(os.path.join(HERE, "golden", "analyzer_digests.json")))
    assert ref["case"] == make_golden.CASE
    got = make_golden.run_case()
    for g, r in zip(got, ref["chunks"]):
        assert g == r, "chunk %d differs from the golden digests" % r["chunk"]
