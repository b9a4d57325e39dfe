"""The FP64 sweep fixture (tests/golden/sweep_fp64_rates.json, written by
tests/golden/make_sweep_fp64.py) holds what the FP64 oracle computes for the sweep points
that spike_amd.sweep samples: the points' parameters are the sampler's, and one entry is
recomputed here and must match exactly."""
import json
import os

import numpy as np

from spike_amd import sweep
from sweep_replay import oracle_rate

HERE = os.path.dirname(os.path.abspath(__file__))


def test_fixture_points_are_the_sampled_points():
    fx = json.load(open(os.path.join(HERE, "golden", "sweep_fp64_rates.json")))
    pts = sweep.sample_points(fx["points_in_sweep"])
    assert fx["realisations"] >= 16 and len(fx["points"]) >= 32
    for e in fx["points"]:
        p = pts[e["index"]]
        assert (e["n"], e["delay_steps"], e["seed"]) == (p.n, p.delay_steps, p.seed)
        assert np.allclose([e["g"], e["nu_ratio"], e["J"]], [p.g, p.nu_ratio, p.J], rtol=0, atol=0)
        assert e["drive_seeds"] == [sweep.drive_seed(p, r) for r in range(fx["realisations"])]
        assert len(e["rate_fp64"]) == fx["realisations"] and min(e["rate_fp64"]) > 0


def test_fixture_entry_recomputed():
    fx = json.load(open(os.path.join(HERE, "golden", "sweep_fp64_rates.json")))
    pts = sweep.sample_points(fx["points_in_sweep"])
    e = min(fx["points"], key=lambda x: x["n"])      # the smallest network: about a second
    r = 3
    rate, _ = oracle_rate(pts[e["index"]], fx["seconds"], precision=64, threads=8,
                          drive_seed=e["drive_seeds"][r])
    assert rate == e["rate_fp64"][r]
