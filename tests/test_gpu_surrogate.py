"""Surrogate screening of a Brunel sweep: what the trained ensemble gains over sweeping
without it, with the batched simulator as the judge.

Training data: 4096 sweep points simulated for 0.2 s each by the batched simulator.
Candidates: 512 new points, all simulated as well so that every claim is checked.

1. Against simulating every candidate: screening simulates under a third of them (the
   uncertain points and the spot checks) and spends under a third of the device time,
   and at least 85 % of the points it answers from the surrogate are within 5 % of the
   simulated rate.
2. Against interpolating the same training data without a network (inverse-distance
   k nearest neighbours in the same feature space): the ensemble's median rate error is
   lower.
3. Fallback: an untrained ensemble fails its spot checks, every point is then simulated,
   and the answers equal the simulator's bit for bit.
"""
import numpy as np
import pytest

from spike_amd import sweep, surrogate as S

pytestmark = pytest.mark.gpu
SECONDS = 0.2


@pytest.fixture(scope="module")
def sweep_data():
    print("[surrogate] simulating 4096 training points and 512 candidates", flush=True)
    train = S.simulate(sweep.sample_points(4096, seed=1000), SECONDS)
    cand_pts = sweep.sample_points(512, seed=900000)
    truth = S.simulate(cand_pts, SECONDS)
    return train, cand_pts, truth


def rate_err(pred, truth):
    return np.abs(pred - truth) / np.maximum(truth, S.RATE_FLOOR_HZ)


def knn_rates(train, pts, k=5):
    xt = S.features(train.points).astype(np.float64)
    xq = S.features(pts).astype(np.float64)
    lt = np.log(train.rate + S.RATE_FLOOR_HZ)
    d = np.sqrt(((xq[:, None, :] - xt[None, :, :]) ** 2).sum(-1))
    nn = np.argsort(d, axis=1)[:, :k]
    w = 1.0 / np.maximum(np.take_along_axis(d, nn, 1), 1e-9)
    return np.exp((w * lt[nn]).sum(1) / w.sum(1)) - S.RATE_FLOOR_HZ


def test_screening_gain_over_simulating_everything_and_over_interpolation(sweep_data):
    train, cand, truth = sweep_data
    sur = S.Surrogate(members=8, batch=1024)
    info = sur.fit(train, steps=3000)
    print(f"[surrogate] trained: {info}", flush=True)
    rate, _, _, spread = sur.predict(cand)
    err = rate_err(rate, truth.rate)
    knn_err = rate_err(knn_rates(train, cand), truth.rate)
    print(f"ensemble median rate error {np.median(err):.4f}, within 5 %: {np.mean(err < 0.05):.3f}; "
          f"kNN median {np.median(knn_err):.4f}, within 5 %: {np.mean(knn_err < 0.05):.3f}", flush=True)
    assert np.median(err) < np.median(knn_err)

    res = S.screen(sur, cand, SECONDS)
    acc = ~res.simulated
    final_err = rate_err(res.rate, truth.rate)
    print(f"screen: simulated {res.simulated.sum()} of {len(cand)} (spread {np.sum(res.reason == 'spread')}, "
          f"spot {np.sum(res.reason == 'spot')}), spot failures {res.spot_failures}, fallback {res.fell_back}; "
          f"device ms {res.sim_device_ms:.1f} vs {truth.device_ms:.1f} for all; "
          f"accepted within 5 %: {np.mean(final_err[acc] < 0.05):.3f}", flush=True)
    assert not res.fell_back
    assert res.simulated.sum() <= len(cand) // 3
    assert res.sim_device_ms <= truth.device_ms / 3
    assert np.mean(final_err[acc] < 0.05) >= 0.85
    # simulated points carry the simulator's values exactly
    assert np.array_equal(res.rate[res.simulated], truth.rate[res.simulated])
    sur.close()


def test_untrained_ensemble_falls_back_to_the_simulator(sweep_data):
    train, cand, truth = sweep_data
    sub = cand[:128]
    sur = S.Surrogate(members=4, batch=128)
    sur.fit(train, steps=0)                 # initial weights only
    res = S.screen(sur, sub, SECONDS, spread_tol=10.0)
    print(f"[surrogate] untrained: spot failures {res.spot_failures}, fallback {res.fell_back}", flush=True)
    assert res.fell_back and res.simulated.all()
    assert np.array_equal(res.rate, truth.rate[:128])
    assert np.array_equal(res.cv, truth.cv[:128]) and np.array_equal(res.sync, truth.sync[:128])
    sur.close()
