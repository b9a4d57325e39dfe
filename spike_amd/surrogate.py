"""Surrogate screening of Brunel parameter sweeps.

An ensemble of MLPs (spike_amd.mlp, trained on the GPU by the library's own kernels) maps a
sweep point (N, g, nu_ext/nu_thr, J, delay) to the statistics the simulator reports (rate,
ISI CV, synchrony).  Training data comes from the batched simulator.  When a sweep is
screened, the simulator stays the fallback and the judge:
  * a point on which the members disagree (ensemble spread of the predicted rate above
    `spread_tol`) is simulated;
  * of the confident points a random fraction (`spot_fraction`, at least `spot_min`) is
    simulated as a spot check; a spot-checked point whose simulated rate is more than
    `rate_tol` off the prediction keeps the simulated value, and if more than
    `max_fail_fraction` of the spot checks fail, every remaining point is simulated;
  * every other confident point takes the ensemble-mean prediction.
"""
from __future__ import annotations

import dataclasses
import math
import time
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import mlp, sweep

# Feature ranges of the sweep (sweep.sample_points defaults); features are scaled to [-1, 1].
RANGES = dict(logn=(math.log(1250.0), math.log(12500.0)), g=(3.0, 8.0), nu=(0.9, 4.0), j_mv=(0.05, 0.2),
              delay=(10.0, 20.0))
RATE_FLOOR_HZ = 0.1   # log(rate + floor): rates near zero do not dominate the loss
SYNC_FLOOR = 1e-3


def features(points: Sequence[sweep.SweepPoint]) -> np.ndarray:
    raw = np.array([[math.log(p.n), p.g, p.nu_ratio, p.J * 1e3, p.delay_steps] for p in points], np.float64)
    lo = np.array([v[0] for v in RANGES.values()])
    hi = np.array([v[1] for v in RANGES.values()])
    return (2.0 * (raw - lo) / (hi - lo) - 1.0).astype(np.float32)


def to_targets(rate, cv, sync) -> np.ndarray:
    return np.stack([np.log(np.asarray(rate) + RATE_FLOOR_HZ), np.asarray(cv),
                     np.log(np.asarray(sync) + SYNC_FLOOR)], 1).astype(np.float64)


def from_targets(t: np.ndarray):
    return np.exp(t[..., 0]) - RATE_FLOOR_HZ, t[..., 1], np.exp(t[..., 2]) - SYNC_FLOOR


@dataclasses.dataclass
class SimData:
    points: List[sweep.SweepPoint]
    rate: np.ndarray
    cv: np.ndarray
    sync: np.ndarray
    device_ms: float
    wall_s: float


def simulate(points: Sequence[sweep.SweepPoint], seconds: float, chunk: int = 256,
             neurons_per_wg: int = 384, log=None) -> SimData:
    """The batched simulator over `points`, `chunk` networks per persistent launch."""
    t0 = time.perf_counter()
    rate, cv, sync, ms = [], [], [], 0.0
    for i in range(0, len(points), chunk):
        pts = list(points[i:i + chunk])
        r = sweep.run_batched(pts, seconds, sweep.default_plan(pts, neurons_per_wg))
        rate.append(r.rate)
        cv.append(r.cv)
        sync.append(r.sync)
        ms += r.ms
        if log:
            log(f"simulated {min(i + chunk, len(points))}/{len(points)} points, device {ms:.1f} ms, "
                f"wall {time.perf_counter() - t0:.1f} s")
    cat = lambda v: np.concatenate(v) if v else np.zeros(0)
    return SimData(list(points), cat(rate), cat(cv), cat(sync), ms, time.perf_counter() - t0)


class Surrogate:
    """Ensemble of `members` MLPs on standardised log-rate, CV and log-synchrony."""

    def __init__(self, members: int = 8, batch: int = 1024, seed: int = 2025):
        self.members, self.batch, self.seed = members, batch, seed
        self.net = mlp.Mlp(members, batch)
        self.mu = np.zeros(3)
        self.sd = np.ones(3)
        self.history: Optional[np.ndarray] = None

    def close(self):
        self.net.close()

    def fit(self, data: SimData, steps: int = 3000, lr0: float = 2e-3, lr_min: float = 1e-5,
            use_graph: bool = True) -> Dict:
        x = features(data.points)
        t = to_targets(data.rate, data.cv, data.sync)
        self.mu, self.sd = t.mean(0), np.maximum(t.std(0), 1e-6)
        tn = ((t - self.mu) / self.sd).astype(np.float32)
        self.net.set_params(mlp.init_params(self.seed, self.members))
        self.net.set_data(x, tn)
        t0 = time.perf_counter()
        self.net.train(steps, self.batch, lr0, lr_min, max(steps, 1), mode=mlp.BATCH_SAMPLED, seed=self.seed,
                       use_graph=use_graph, record_loss=False)
        tm = self.net.timing()
        wall = time.perf_counter() - t0
        rows = (len(x) // mlp.ROW_TILE) * mlp.ROW_TILE
        loss = None
        if rows >= mlp.ROW_TILE:
            loss, _ = self.net.gradients(min(rows, self.batch))
        return dict(steps=steps, device_ms=tm.total_ms, wall_s=wall,
                    final_loss=None if loss is None else loss.tolist())

    def predict(self, points: Sequence[sweep.SweepPoint]):
        """Ensemble mean of (rate, cv, sync) and the members' spread of log-rate."""
        y = self.net.predict(features(points)).astype(np.float64)     # [members][n][3]
        t = y * self.sd + self.mu
        mean = t.mean(0)
        rate, cv, sync = from_targets(mean)
        spread = t[:, :, 0].std(0)   # spread of log(rate + floor): a relative rate spread
        return rate, cv, sync, spread


@dataclasses.dataclass
class ScreenResult:
    rate: np.ndarray
    cv: np.ndarray
    sync: np.ndarray
    simulated: np.ndarray          # bool per point
    reason: np.ndarray             # "spread", "spot", "fallback" or "" (surrogate value)
    spot_failures: int
    fell_back: bool
    sim_device_ms: float
    predict_ms: float


def screen(sur: Surrogate, points: Sequence[sweep.SweepPoint], seconds: float, spread_tol: float = 0.05,
           spot_fraction: float = 0.1, spot_min: int = 8, rate_tol: float = 0.05,
           max_fail_fraction: float = 0.05, seed: int = 7, log=None) -> ScreenResult:
    n = len(points)
    rate, cv, sync, spread = sur.predict(points)
    predict_ms = sur.net.timing().forward_ms
    reason = np.array([""] * n, dtype=object)
    unc = spread > spread_tol
    reason[unc] = "spread"
    conf = np.nonzero(~unc)[0]
    rng = np.random.default_rng(seed)
    k = min(len(conf), max(spot_min, int(math.ceil(spot_fraction * len(conf)))))
    spot = rng.choice(conf, size=k, replace=False) if k else np.zeros(0, int)
    reason[spot] = "spot"
    first = np.nonzero(reason != "")[0]
    sim = simulate([points[i] for i in first], seconds) if len(first) else None
    out_rate, out_cv, out_sync = rate.copy(), cv.copy(), sync.copy()
    simulated = np.zeros(n, bool)
    ms = 0.0
    fails = 0
    if sim is not None:
        ms += sim.device_ms
        for j, i in enumerate(first):
            if reason[i] == "spot" and abs(rate[i] - sim.rate[j]) > rate_tol * max(sim.rate[j], RATE_FLOOR_HZ):
                fails += 1
            out_rate[i], out_cv[i], out_sync[i] = sim.rate[j], sim.cv[j], sim.sync[j]
            simulated[i] = True
    fell_back = k > 0 and fails > max_fail_fraction * k
    if fell_back:
        rest = np.nonzero(~simulated)[0]
        if log:
            log(f"{fails}/{k} spot checks off by more than {rate_tol:.0%}: simulating the other {len(rest)} points")
        sim2 = simulate([points[i] for i in rest], seconds)
        ms += sim2.device_ms
        for j, i in enumerate(rest):
            out_rate[i], out_cv[i], out_sync[i] = sim2.rate[j], sim2.cv[j], sim2.sync[j]
            simulated[i] = True
            reason[i] = "fallback"
    return ScreenResult(out_rate, out_cv, out_sync, simulated, reason, fails, fell_back, ms, predict_ms)


def main(argv=None):
    """Train the surrogate on a simulated sweep and measure what screening gains:
    python -m spike_amd.surrogate --train-points 50000 --seconds 1.0 --test-points 2048"""
    import argparse
    import json
    ap = argparse.ArgumentParser()
    ap.add_argument("--train-points", type=int, default=50000)
    ap.add_argument("--test-points", type=int, default=2048)
    ap.add_argument("--seconds", type=float, default=1.0)
    ap.add_argument("--members", type=int, default=8)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--steps", type=int, default=8000)
    ap.add_argument("--out", default="")
    a = ap.parse_args(argv)
    log = lambda m: print(m, flush=True)
    train = simulate(sweep.sample_points(a.train_points, seed=1000), a.seconds, log=log)
    test_pts = sweep.sample_points(a.test_points, seed=900000)
    truth = simulate(test_pts, a.seconds)
    sur = Surrogate(members=a.members, batch=a.batch)
    info = sur.fit(train, steps=a.steps)
    rate, cv, sync, spread = sur.predict(test_pts)
    pred_ms = sur.net.timing().forward_ms
    err = np.abs(rate - truth.rate) / np.maximum(truth.rate, RATE_FLOOR_HZ)
    res = screen(sur, test_pts, a.seconds, log=log)
    fin = np.abs(res.rate - truth.rate) / np.maximum(truth.rate, RATE_FLOOR_HZ)
    acc = ~res.simulated
    out = dict(train_points=a.train_points, test_points=a.test_points, seconds=a.seconds, members=a.members,
               train_sim_device_ms=train.device_ms, train_sim_wall_s=train.wall_s, fit=info,
               test_all_sim_device_ms=truth.device_ms, predict_ms=pred_ms,
               rate_err_median=float(np.median(err)), rate_err_p90=float(np.percentile(err, 90)),
               within_5pc=float(np.mean(err < 0.05)),
               cv_abs_err_median=float(np.median(np.abs(cv - truth.cv))),
               screen=dict(simulated=int(res.simulated.sum()), spread=int(np.sum(res.reason == "spread")),
                           spot=int(np.sum(res.reason == "spot")), spot_failures=res.spot_failures,
                           fell_back=bool(res.fell_back), sim_device_ms=res.sim_device_ms,
                           accepted_within_5pc=float(np.mean(fin[acc] < 0.05)) if acc.any() else None,
                           device_time_fraction=res.sim_device_ms / truth.device_ms))
    print(json.dumps(out), flush=True)
    if a.out:
        with open(a.out, "w") as f:
            json.dump(out, f, indent=1)
    sur.close()


if __name__ == "__main__":
    main()
