"""The FP64 COBA statistics fixture (tests/golden/coba_vogels_fp64_2s.json) holds what the
FP64 oracle computes: one seed is recomputed here and must match exactly."""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_coba_fp64  # noqa: E402


def test_fixture_entry_recomputed():
    with open(os.path.join(HERE, "golden", "coba_vogels_fp64_2s.json")) as f:
        gold = json.load(f)
    assert sorted(int(s) for s in gold["seeds"]) == make_coba_fp64.SEEDS
    e = json.loads(json.dumps(make_coba_fp64.fp64_entry(45, threads=8, seconds=2.0)))
    assert e == gold["seeds"]["45"]
