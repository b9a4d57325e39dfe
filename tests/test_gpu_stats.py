"""FP64 spike statistics on the GPU (rate, ISI CV, synchrony).

1. Same run, same raster: the device statistics equal float64 statistics computed here
   from the FP32 oracle's raster (which is bit-identical to the GPU's), on both the
   persistent and the per-phase path, with a window that starts late and one that ends
   early.
2. FP32 GPU against the FP64 oracle on longer chaotic runs: population rate, ISI CV and the
   final weight histogram of a plastic network agree within 2 %.
"""
import numpy as np
import pytest

import oracle_py
from spike_amd import abi, builders
from spike_amd.netdesc import Rule, RULE_WD

pytestmark = pytest.mark.gpu


def expected_stats(ids, times, n, dt, t0, steps, bin_steps):
    """float64 statistics of a raster over the window [t0, t0 + steps)."""
    dt = float(np.float32(dt))   # the network's timestep is a float32
    st = np.rint(times.astype(np.float64) / dt).astype(np.int64)
    keep = (st >= t0) & (st < t0 + steps)
    ids, times, st = ids[keep], times[keep], st[keep]
    spikes = len(ids)
    cvs = []
    for i in range(n):
        t = np.sort(times[ids == i].astype(np.float64))
        if len(t) >= 3:
            isi = np.diff(t)
            m = isi.mean()
            cvs.append(np.sqrt(max(0.0, (isi * isi).mean() - m * m)) / m)
    nb = (steps + bin_steps - 1) // bin_steps
    hist = np.bincount((st - t0) // bin_steps, minlength=nb)[:nb].astype(np.float64)
    mean = hist.mean()
    return dict(spikes=spikes, rate=spikes / (n * steps * dt), cv=float(np.mean(cvs)) if cvs else 0.0,
                ncv=len(cvs), sync=float(hist.var() / mean) if mean > 0 else 0.0, bins=nb)


@pytest.mark.parametrize("launches", [False, True])
@pytest.mark.parametrize("window", ["late_start", "early_end"])
def test_device_stats_equal_raster_stats(launches, window):
    d = builders.brunel(n_exc=800, n_inh=200, seed=11, monitors=(0,))
    G = d.timestep_grouping()
    print(f"[stats] brunel N=1000 launches={launches} window={window}", flush=True)
    gpu = abi.Network(d)
    ora = oracle_py.OracleNetwork(d, precision=32)
    pre_groups = 20 if window == "late_start" else 0
    gpu.run(0.0, pre_groups, force_launches=launches)
    ora.run(0.0, pre_groups)
    start = float(d.group_time(0.0, pre_groups))
    ng = 200
    max_steps = ng * G if window == "late_start" else 1234
    bin_steps = 10
    gpu.stats_begin(bin_steps, max_steps)
    gpu.run(start, ng, force_launches=launches)
    ora.run(start, ng)
    s = gpu.stats()
    oi, ot = ora.take_spikes(0)
    t0 = pre_groups * G
    steps = min(ng * G, max_steps)
    e = expected_stats(oi, ot, d.n_neurons, d.dt, t0, steps, bin_steps)
    assert s.steps == steps and s.bins == e["bins"]
    assert s.spikes == e["spikes"] > 0
    assert s.cv_neurons == e["ncv"]
    assert abs(s.rate_hz - e["rate"]) <= 1e-12 * e["rate"]
    assert abs(s.isi_cv - e["cv"]) <= 1e-9 * e["cv"]
    assert abs(s.synchrony - e["sync"]) <= 1e-9 * e["sync"]


def test_step_group_stats_match_run_stats():
    """Statistics accumulate the same way on the per-group entry point."""
    d = builders.brunel(n_exc=400, n_inh=100, seed=12)
    G = d.timestep_grouping()
    a = abi.Network(d)
    b = abi.Network(d)
    a.stats_begin(5, 100 * G)
    b.stats_begin(5, 100 * G)
    a.run(0.0, 100)
    for s in range(100):
        b.step_group(d.group_time(0.0, s))
    sa, sb = a.stats(), b.stats()
    assert (sa.spikes, sa.cv_neurons, sa.bins) == (sb.spikes, sb.cv_neurons, sb.bins)
    assert sa.isi_cv == sb.isi_cv and sa.synchrony == sb.synchrony and sa.rate_hz == sb.rate_hz


def test_fp32_gpu_statistics_within_2pc_of_fp64_oracle():
    """Chaotic run: the FP32 trajectory departs from the FP64 one, the statistics do not."""
    d = builders.brunel(n_exc=1600, n_inh=400, seed=5, monitors=(0,))
    ng = d.groups_for(2.0)
    G = d.timestep_grouping()
    print(f"[stats] brunel N=2000 2 s FP32 GPU vs FP64 oracle", flush=True)
    gpu = abi.Network(d)
    gpu.stats_begin(10, ng * G)
    gpu.run(0.0, ng)
    s = gpu.stats()
    o = oracle_py.OracleNetwork(d, precision=64, threads=8)
    o.run(0.0, ng)
    oi, ot = o.take_spikes(0)
    e = expected_stats(oi, ot, d.n_neurons, d.dt, 0, ng * G, 10)
    print(f"GPU FP32 rate {s.rate_hz:.4f} cv {s.isi_cv:.4f} sync {s.synchrony:.3f}; "
          f"FP64 rate {e['rate']:.4f} cv {e['cv']:.4f} sync {e['sync']:.3f}", flush=True)
    assert abs(s.rate_hz - e["rate"]) <= 0.02 * e["rate"]
    assert abs(s.isi_cv - e["cv"]) <= 0.02 * e["cv"]
    assert abs(s.synchrony - e["sync"]) <= 0.15 * e["sync"]


def test_fp32_plastic_weight_histogram_within_2pc_of_fp64():
    """WD STDP on E->E of a 4000-neuron network for 2 s at a learning rate where each weight
    averages many pairings: the FP32 weight distribution (mean, spread, 20-bin histogram)
    stays within 2 % of the FP64 one although the spike trains have decorrelated.  (At
    1000 neurons the weights of synapses sharing a neuron are too correlated for a 2 %
    histogram comparison: there two FP64 runs that differ only in the random seed already
    differ by about 6 %.)"""
    d = builders.brunel(n_exc=3200, n_inh=800, seed=6)
    ne = 3200
    exc = np.nonzero((d.pre >= 0) & (d.pre < ne) & (d.post < ne))[0].astype(np.int64)
    d.rules.append(Rule(RULE_WD, exc, builders.wd_params(lam=0.05, alpha=2.02, w_max=0.3e-3)))
    ng = d.groups_for(2.0)
    print(f"[stats] plastic brunel N=4000 2 s FP32 GPU vs FP64 oracle weights", flush=True)
    gpu = abi.Network(d)
    gpu.run(0.0, ng)
    wg = gpu.weights()[exc].astype(np.float64)
    o = oracle_py.OracleNetwork(d, precision=64, threads=16)
    o.run(0.0, ng)
    wo = o.weights()[exc]
    edges = np.linspace(min(wg.min(), wo.min()), max(wg.max(), wo.max()) * (1 + 1e-9), 21)
    hg, _ = np.histogram(wg, bins=edges)
    ho, _ = np.histogram(wo, bins=edges)
    l1 = np.abs(hg - ho).sum() / (2.0 * len(exc))
    print(f"weight histogram L1/2 = {l1:.4f}; mean GPU {wg.mean():.4e} FP64 {wo.mean():.4e}; "
          f"std GPU {wg.std():.4e} FP64 {wo.std():.4e}", flush=True)
    assert wo.std() > 0.02 * wo.mean(), "weights did not spread"
    assert abs(wg.mean() - wo.mean()) <= 0.02 * wo.mean()
    assert abs(wg.std() - wo.std()) <= 0.02 * wo.std()
    assert l1 <= 0.02
