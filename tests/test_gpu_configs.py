"""The benchmark configurations at their real sizes on the GPU.

1. The first 1000 timesteps are bit-exact against the FP32 oracle (spike rasters, membrane
   potentials, last-spike times, spike history, the input ring, conductances, weights and
   plasticity traces) for:
   * Brunel10K, config form (8000 E + 2000 I, 10^7 synapses, per-neuron Poisson drive) and
     reference form (+10^4 Poisson inputs, +10^7 synapses);
   * the plain small network (1000 Poisson inputs -> 1000 conductance LIF, p = 0.1; with
     U(0, 1) nS weights and 20 Hz inputs its LIF neurons stay below threshold, so this case
     checks the inputs, conductances and membranes);
   * the COBA network with Vogels iSTDP on I->E (3200 E + 800 I, p = 0.02), plasticity on;
   * the large plastic network at 1/100 of its size (5000 neurons x 100 outgoing synapses,
     delays 1..20 steps so G = 1, nearest-spike WD STDP on E->E, 250 Poisson inputs).
2. 2 s COBA + iSTDP runs on the GPU, with FP32 or FP16 weight storage, stay within 2 % of
   the FP64 oracle (BF16: the stated tolerances of COBA_TOL) in population rate, ISI CV
   and the I->E weight distribution (mean and 20-bin histogram), as means over 8
   connectivity seeds: single runs decorrelate within milliseconds, and the FP64 oracle's
   own rate and CV move by 1-3 % per seed when its resting potentials move by one ulp.
3. 2 s runs of the large plastic configuration at 1/25 of its size (full-size in-degrees,
   drive and rule), FP32 and FP16 storage, stay within 2 % of the FP64 oracle in rate, ISI
   CV, mean E->E weight and the E->E weight histogram, as means over 4 seeds.
"""
import json
import os

import numpy as np
import pytest

import oracle_py
from chaos_stats import hist_l1, histogram, isi_cv
from spike_amd import abi, builders

pytestmark = pytest.mark.gpu

from test_gpu_parity import _compare  # noqa: E402


def _bit_exact_first_steps(d, steps=1000, tag="", force_launches=False):
    G = d.timestep_grouping()
    ng = -(-steps // G)
    print(f"[configs] {tag}: N={d.n_neurons} Nin={d.n_inputs} S={d.n_synapses} G={G} groups={ng}", flush=True)
    gpu = abi.Network(d)
    ora = oracle_py.OracleNetwork(d, precision=32, threads=16)
    gi, oi = gpu.info(), ora.info()
    assert (gi.timestep_grouping, gi.ring_slots, gi.history_words, gi.frac_bits) == \
        (oi["G"], oi["B"], oi["W"], oi["frac"])
    gpu.run(0.0, ng, force_launches=force_launches)
    ora.run(0.0, ng)
    c, oc = gpu.counters(), ora.counters()
    assert c.spikes_neurons == oc["spikes_neurons"], tag
    assert c.spikes_inputs == oc["spikes_inputs"], tag
    assert c.spikes_neurons + c.spikes_inputs > 0, tag
    print(f"  fused={gi.fused} record_bytes={gi.record_bytes}: {c.spikes_neurons} spikes, "
          f"{c.synaptic_events} deliveries", flush=True)
    _compare(gpu, ora, d, tag)
    gpu.close()
    ora.close()


@pytest.mark.parametrize("form", ["config", "reference"])
def test_brunel10k_first_1000_steps(form):
    _bit_exact_first_steps(builders.brunel(seed=42, form=form, monitors=(0,)), tag=f"brunel10k {form}")


def test_plain_small_first_1000_steps():
    _bit_exact_first_steps(builders.plain_small(seed=42, monitors=(0, 1)), tag="plain small")


def test_coba_vogels_first_1000_steps():
    _bit_exact_first_steps(builders.coba_vogels(seed=42, monitors=(0,)), tag="coba + vogels")


def test_large_plastic_scaled_first_1000_steps():
    d = builders.large_plastic(n=5000, out_degree=100, monitors=(0, 1), generated=False)
    _bit_exact_first_steps(d, tag="large plastic / 100")


@pytest.mark.parametrize("launches", [False, True])
def test_large_plastic_threaded_preparation(launches):
    """Above 2^20 synapses the host preparation runs on threads (validation, counting sorts,
    gathers, the rule's index sorts); the device layout and so every result must be the
    sequential one: bit-exact against the oracle with FP16 weights, on the persistent
    kernel and on the per-phase kernels (whose rule is driven by the spike-indexed arrays)."""
    d = builders.large_plastic(n=10000, out_degree=170, monitors=(0, 1), generated=False)
    d.weight_format = 2
    assert len(d.rules[0].synapses) > (1 << 20)
    _bit_exact_first_steps(d, steps=300, tag=f"large plastic / 170, threaded preparation, launches={launches}",
                           force_launches=launches)


GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "coba_vogels_fp64_2s.json")

# Relative tolerances against FP64 per weight format, on means over the fixture's seeds:
# (rate, ISI CV, mean I->E weight, mean histogram L1/2).  FP32 and FP16 storage meet the
# 2 % criterion.  BF16 is stated, not held to it: its 8-bit significand moves the fixed
# weights themselves (0.06 by +0.1 %, 0.67 by -0.3 %) and biases the mean plastic weight
# (+0.25 %), and its spacing near 0.67 (0.0039) is a third of the weights' spread, so its
# histogram is compared with that of the FP64 weights rounded to bfloat16.
COBA_TOL = {0: (0.02, 0.02, 0.02, 0.02), 1: (0.03, 0.03, 0.005, 0.10), 2: (0.02, 0.02, 0.02, 0.02)}


def _coba_ensemble(gold_path, fmt):
    with open(gold_path) as f:
        gold = json.load(f)
    rows = []
    for seed, g in gold["seeds"].items():
        d = builders.coba_vogels(seed=int(seed), monitors=(0,))
        d.weight_format = fmt
        print(f"[configs] coba + vogels {gold['seconds']:g} s seed {seed} weights format {fmt}", flush=True)
        gpu = abi.Network(d)
        gpu.run(0.0, d.groups_for(gold["seconds"]))
        ids, times = gpu.take_spikes(0)
        w = gpu.weights()[d.rules[0].synapses].astype(np.float64)
        gpu.close()
        hist = histogram(w, np.asarray(g["edges"]))
        l1 = hist_l1(hist, g["hist_bf16"] if fmt == 1 else g["hist"])
        rows.append((len(ids), isi_cv(ids, times, d.n_neurons), w.mean(), l1, g["spikes"], g["cv"], g["w_mean"]))
        r = rows[-1]
        print(f"  spikes {r[0]} vs {r[4]} ({(r[0] - r[4]) / r[4]:+.2%}), CV {r[1]:.4f} vs {r[5]:.4f} "
              f"({(r[1] - r[5]) / r[5]:+.2%}), mean weight {(r[2] - r[6]) / r[6]:+.3%}, histogram L1/2 {l1:.4f}",
              flush=True)
    a = np.array(rows)
    rate, cv, wm = [(a[:, k].mean() - a[:, k + 4].mean()) / a[:, k + 4].mean() for k in range(3)]
    l1 = a[:, 3].mean()
    print(f"means over {len(a)} seeds: rate {rate:+.2%}, CV {cv:+.2%}, mean weight {wm:+.3%}, "
          f"histogram L1/2 {l1:.4f}", flush=True)
    return rate, cv, wm, l1


@pytest.mark.parametrize("fmt", [0, 1, 2])
def test_coba_vogels_2s_ensemble_within_2pc_of_fp64(fmt):
    """COBA + iSTDP, 2 s on the GPU with FP32, BF16 (Kahan) or FP16 (Kahan) weight storage
    against FP64 oracle statistics of the same networks (tests/golden/make_coba_fp64.py),
    over 8 connectivity seeds: a single 2 s run is one realisation of a chaotic network
    whose rate and ISI CV move by 1-3 % under a one-ulp perturbation (chaos_stats), so the
    criterion is on the means over the seeds; per-seed differences are printed."""
    rate, cv, wm, l1 = _coba_ensemble(GOLD, fmt)
    t_rate, t_cv, t_mean, t_hist = COBA_TOL[fmt]
    assert abs(rate) <= t_rate
    assert abs(cv) <= t_cv
    assert abs(wm) <= t_mean
    assert l1 <= t_hist


GOLD10 = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "coba_vogels_fp64_10s.json")


@pytest.mark.parametrize("fmt", [0, 1, 2])
def test_coba_vogels_10s_ensemble_within_2pc_of_fp64(fmt):
    """The COBA configuration's own duration, 10 s (tests/golden/make_coba_fp64.py 8 10):
    the same comparison and tolerances as the 2 s runs, over the same 8 seeds."""
    rate, cv, wm, l1 = _coba_ensemble(GOLD10, fmt)
    t_rate, t_cv, t_mean, t_hist = COBA_TOL[fmt]
    assert abs(rate) <= t_rate
    assert abs(cv) <= t_cv
    assert abs(wm) <= t_mean
    assert l1 <= t_hist


GOLD_LARGE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "large_plastic_fp64_2s.json")


@pytest.mark.parametrize("fmt", [0, 2])
def test_large_plastic_2s_ensemble_within_2pc_of_fp64(fmt):
    """The large plastic configuration at 1/25 of its size (2e4 neurons with the full-size
    in-degrees, drive and rule; tests/golden/make_large_fp64.py), 2 s on the GPU with FP32
    or FP16 (Kahan) weight storage against the FP64 oracle over 4 seeds: population rate,
    ISI CV, mean E->E weight and the E->E weight histogram over [0, 1], as means over the
    seeds (single runs of a chaotic network are single realisations)."""
    with open(GOLD_LARGE) as f:
        gold = json.load(f)
    rows = []
    for seed, g in gold["seeds"].items():
        d = builders.large_plastic(n=gold["n_neurons"], seed=int(seed), monitors=(0,))
        d.weight_format = fmt
        print(f"[configs] large plastic / 25, 2 s, seed {seed}, weights format {fmt}", flush=True)
        gpu = abi.Network(d)
        gpu.run(0.0, d.groups_for(gold["seconds"]))
        ids, times = gpu.take_spikes(0)
        w = gpu.weights()
        gpu.close()
        pre, post, _, _, _ = oracle_py.outdegree_connectivity(d.seed, d.n_neurons, d.n_inputs, d.outdegree, threads=8)
        ee = (pre >= 0) & (pre < d.outdegree.n_exc) & (post < d.outdegree.n_exc)
        w = w[ee].astype(np.float64)
        l1 = hist_l1(histogram(w, np.asarray(g["edges"])), g["hist"])
        rows.append((len(ids), isi_cv(ids, times, d.n_neurons), w.mean(), l1, g["spikes"], g["cv"], g["w_mean"]))
        r = rows[-1]
        print(f"  spikes {r[0]} vs {r[4]} ({(r[0] - r[4]) / r[4]:+.2%}), CV {r[1]:.4f} vs {r[5]:.4f} "
              f"({(r[1] - r[5]) / r[5]:+.2%}), mean weight {(r[2] - r[6]) / r[6]:+.3%}, histogram L1/2 {l1:.4f}",
              flush=True)
    a = np.array(rows)
    rate, cv, wm = [(a[:, k].mean() - a[:, k + 4].mean()) / a[:, k + 4].mean() for k in range(3)]
    l1 = a[:, 3].mean()
    print(f"means over {len(a)} seeds: rate {rate:+.2%}, CV {cv:+.2%}, mean weight {wm:+.3%}, "
          f"histogram L1/2 {l1:.4f}", flush=True)
    assert abs(rate) <= 0.02
    assert abs(cv) <= 0.02
    assert abs(wm) <= 0.02
    assert l1 <= 0.02
