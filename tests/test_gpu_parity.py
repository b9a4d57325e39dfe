"""Bit-exact parity of the HIP backend against the FP32 CPU oracle.

Both run the same NetDesc; spike rasters (id, time), membrane potentials, the input
ring, spike history, conductances, weights and plasticity traces must be identical.
"""
import numpy as np
import pytest

import oracle_py
from spike_amd import abi, builders
from spike_amd.netdesc import (SYN_CONDUCTANCE, Rule, RULE_WD, RULE_VOGELS, RULE_EVANS, RULE_CUSTOM, WEIGHTS_BF16,
                               WEIGHTS_FP16, WEIGHTS_FP32)

pytestmark = pytest.mark.gpu


def _compare(gpu, ora, d, tag=""):
    for m in range(len(d.monitors)):
        gi, gt = gpu.take_spikes(m)
        oi, ot = ora.take_spikes(m)
        assert len(gi) == len(oi), f"{tag} monitor {m}: {len(gi)} vs {len(oi)} spikes"
        assert np.array_equal(gi, oi), f"{tag} monitor {m} ids differ"
        assert np.array_equal(gt, ot), f"{tag} monitor {m} times differ"
    assert np.array_equal(gpu.membrane().astype(np.float64), ora.membrane()), f"{tag} V differs"
    assert np.array_equal(gpu.last_spike(0), ora.last_spike(0)), f"{tag} last spike differs"
    if d.n_inputs:
        assert np.array_equal(gpu.last_spike(1), ora.last_spike(1)), f"{tag} input last spike"
        assert np.array_equal(gpu.history(1), ora.history(1)), f"{tag} input history differs"
    assert np.array_equal(gpu.history(0), ora.history(0)), f"{tag} history differs"
    assert np.array_equal(gpu.ring(), ora.ring()), f"{tag} input ring differs"
    if d.kind != 0:   # conductances, or the current traces of current synapses
        assert np.array_equal(gpu.conductances().astype(np.float64), ora.conductances()), f"{tag} g"
    if d.n_synapses:
        assert np.array_equal(gpu.weights().astype(np.float64), ora.weights()), f"{tag} weights"
    for r in range(len(d.rules)):
        ga, gb = gpu.traces(r)
        oa, ob = ora.traces(r)
        assert np.array_equal(ga.astype(np.float64), oa) and np.array_equal(gb.astype(np.float64), ob), \
            f"{tag} traces of rule {r}"


def _run_both(d, ngroups, fused=True, start=0.0, partitions=0, launches=False, expect_fused=None):
    print(f"[parity] {d.name} N={d.n_neurons} Nin={d.n_inputs} S={d.n_synapses} G={d.timestep_grouping()} "
          f"groups={ngroups} fused={fused} P={partitions} launches={launches}", flush=True)
    gpu = abi.Network(d, partitions=partitions)
    if expect_fused is not None:
        assert gpu.info().fused == int(expect_fused), gpu.fused_status()
    ora = oracle_py.OracleNetwork(d, precision=32)
    gi, oi = gpu.info(), ora.info()
    assert (gi.timestep_grouping, gi.ring_slots, gi.history_words, gi.frac_bits) == \
        (oi["G"], oi["B"], oi["W"], oi["frac"])
    if fused:
        gpu.run(start, ngroups, force_launches=launches)
        ora.run(start, ngroups)
    else:
        for s in range(ngroups):
            t = d.group_time(start, s)
            gpu.step_phases(t)
            ora.step_phases(t)
    return gpu, ora


@pytest.mark.parametrize("seed", [1, 2])
@pytest.mark.parametrize("fused", [True, False])
def test_voltage_random(seed, fused):
    d = builders.random_net(seed=seed, input_rate=400.0, whi=4e-3)
    gpu, ora = _run_both(d, 300, fused=fused)
    _compare(gpu, ora, d, f"voltage seed={seed}")


@pytest.mark.parametrize("partitions", [1, 3, 8])
def test_persistent_kernel_partitions(partitions):
    d = builders.random_net(seed=12, input_rate=400.0, whi=4e-3, dmin=2, dmax=7, ext_rate=2000.0,
                            ext_w=1e-3)
    gpu, ora = _run_both(d, 150, partitions=partitions, expect_fused=True)
    assert gpu.info().partitions == partitions
    _compare(gpu, ora, d, f"persistent P={partitions}")


def test_launch_path_through_run():
    d = builders.random_net(seed=13, input_rate=400.0, whi=4e-3, dmin=2, dmax=7)
    gpu, ora = _run_both(d, 150, launches=True)
    _compare(gpu, ora, d, "launch path")


def test_run_in_pieces_matches_one_run():
    """Two spk_run calls continue the state exactly (chunk boundaries and restarts)."""
    d = builders.random_net(seed=14, input_rate=500.0, whi=4e-3, dmin=3, dmax=8)
    gpu = abi.Network(d)
    ora = oracle_py.OracleNetwork(d, precision=32)
    gpu.run(0.0, 70)
    t = float(d.group_time(0.0, 70))
    gpu.run(t, 130, force_launches=True)
    t2 = float(d.group_time(t, 130))
    gpu.run(t2, 1500)
    ora.run(0.0, 70)
    ora.run(t, 130)
    ora.run(t2, 1500)
    _compare(gpu, ora, d, "pieces")


def test_voltage_grouped_external():
    d = builders.random_net(seed=5, dmin=4, dmax=9, input_rate=300.0, ext_rate=3000.0,
                            ext_w=1.5e-3, whi=3e-3)
    assert d.timestep_grouping() == 4
    gpu, ora = _run_both(d, 200)
    _compare(gpu, ora, d, "grouped+external")


@pytest.mark.parametrize("labels", [1, 2, 3])
def test_conductance_random(labels):
    d = builders.random_net(seed=11 + labels, kind=SYN_CONDUCTANCE, n_labels=labels, whi=1.0,
                            scale=lambda rng, S: rng.uniform(0.5e-9, 2e-9, S), input_rate=500.0,
                            dmin=2, dmax=5)
    gpu, ora = _run_both(d, 250)
    _compare(gpu, ora, d, f"conductance L={labels}")


def test_brunel_small_1000_steps():
    d = builders.brunel(n_exc=800, n_inh=200, seed=3, monitors=(0,))
    ng = 1000 // d.timestep_grouping() + 1
    gpu, ora = _run_both(d, ng, expect_fused=True)
    _compare(gpu, ora, d, "brunel small")


def test_brunel_reference_form_small():
    d = builders.brunel(n_exc=400, n_inh=100, seed=4, form="reference", monitors=(0, 1))
    d.rates[:] = 40.0
    gpu, ora = _run_both(d, 40)
    _compare(gpu, ora, d, "brunel reference form")


def test_stimulus_switch_and_reset():
    d = builders.random_net(seed=21, n_stimuli=3, input_rate=600.0)
    gpu = abi.Network(d)
    ora = oracle_py.OracleNetwork(d, precision=32)
    gpu.run(0.0, 50)
    ora.run(0.0, 50)
    _compare(gpu, ora, d, "before switch")
    for lib, h, fn in ((gpu.lib, gpu.h, "spk_select_stimulus"), (ora.lib, ora.h, "or_select_stimulus")):
        assert getattr(lib, fn)(h, 2) == 0
    t = d.group_time(0.0, 50)
    gpu.run(float(t), 50)
    ora.run(float(t), 50)
    _compare(gpu, ora, d, "after switch")
    gpu.reset_state()
    ora.reset_state()
    gpu.run(0.0, 30)
    ora.run(0.0, 30)
    _compare(gpu, ora, d, "after reset")


def test_grouping_change_to_one():
    """VogelsAbbottNet --NOTG: timestep_grouping forced to 1 after finalise."""
    d = builders.random_net(seed=8, dmin=3, dmax=7, input_rate=400.0)
    gpu = abi.Network(d)
    ora = oracle_py.OracleNetwork(d, precision=32)
    gpu.set_grouping(1)
    ora.set_grouping(1)
    gpu.run(0.0, 300)
    ora.run(0.0, 300)
    _compare(gpu, ora, d, "G=1")


def _plastic_net(kind, seed):
    d = builders.random_net(seed=seed, n_neurons=96, n_inputs=48, p=0.25, input_rate=500.0,
                            whi=3e-3, dmin=2, dmax=6)
    S = d.n_synapses
    rng = np.random.default_rng(seed)
    syn = np.sort(rng.choice(S, size=S // 2, replace=False)).astype(np.int64)
    if kind == RULE_WD:
        p = builders.wd_params(lam=1e-2, alpha=2.02, w_max=4e-3)
    elif kind == RULE_CUSTOM:
        p = builders.custom_params(a_minus=0.8, a_plus=1.0, a_star=0.05, w_max=4e-3, learning_rate=2e-5)
    elif kind == RULE_VOGELS:
        p = builders.vogels_params(lr=1e-5, rate=10.0)
    else:
        d.weight = rng.uniform(0.0, 1.0, S).astype(np.float32)
        d.scale = np.full(S, np.float32(3e-3), np.float32)
        p = builders.evans_params()
    d.rules.append(Rule(kind, syn, p))
    return d


@pytest.mark.parametrize("kind", [RULE_WD, RULE_VOGELS, RULE_EVANS, RULE_CUSTOM])
@pytest.mark.parametrize("fused", [True, False])
def test_plasticity(kind, fused):
    d = _plastic_net(kind, 30 + kind)
    gpu, ora = _run_both(d, 200, fused=fused, partitions=5)
    _compare(gpu, ora, d, f"rule {kind}")


@pytest.mark.parametrize("fmt", [WEIGHTS_BF16, WEIGHTS_FP16])
@pytest.mark.parametrize("kind", [RULE_WD, RULE_VOGELS, RULE_EVANS, RULE_CUSTOM])
@pytest.mark.parametrize("fused", [True, False])
def test_plasticity_16bit_weights(kind, fused, fmt):
    """16-bit weights (Kahan-compensated updates) and binary16 traces: bit-exact against the
    FP32 oracle storing the same formats."""
    d = _plastic_net(kind, 50 + kind)
    d.weight_format = fmt
    gpu, ora = _run_both(d, 200, fused=fused, partitions=5)
    _compare(gpu, ora, d, f"rule {kind} weights format {fmt}")


@pytest.mark.parametrize("fmt", [WEIGHTS_FP32, WEIGHTS_FP16])
@pytest.mark.parametrize("nearest", [0, 1])
def test_wd_spike_driven_form_across_paths(nearest, fmt):
    """The per-phase path runs the weight-dependent rule spike-driven (traces per neuron,
    only synapses with a pre- or postsynaptic spike in the group updated) after one
    per-synapse group; the persistent kernel runs it per synapse.  Alternating the two,
    reading traces and weights in between, and setting weights below 0 (which only the
    per-synapse form clamps, so the next group is one) keeps every array bit-exact."""
    d = _plastic_net(RULE_WD, 60 + nearest)
    d.rules[0].params["nearest_spike_only"] = nearest
    d.weight_format = fmt
    print(f"[parity] wd spike-driven across paths nearest={nearest} fmt={fmt}", flush=True)
    gpu = abi.Network(d, partitions=5)
    ora = oracle_py.OracleNetwork(d, precision=32)
    assert gpu.info().fused == 1
    t, tag = 0.0, "wd paths"
    for k, (ng, launches) in enumerate([(40, True), (30, False), (50, True), (20, True)]):
        if k == 3:
            w = gpu.weights()
            w[d.rules[0].synapses[::7]] = np.float32(-1e-3)
            gpu.set_weights(w)
            ora.set_weights(w)
        gpu.run(t, ng, force_launches=launches)
        ora.run(t, ng)
        t = float(d.group_time(t, ng))
        _compare(gpu, ora, d, f"{tag} piece {k}")
    assert gpu.weights()[d.rules[0].synapses].min() >= 0.0


def test_wd_spike_driven_form_csr_order():
    """A network that the persistent kernel cannot hold (one partition whose input-ring
    slice exceeds LDS) keeps its plastic synapses in CSR order, which already groups them
    by presynaptic neuron: the spike-driven kernels then index them without a permutation."""
    d = builders.random_net(seed=71, n_neurons=2000, n_inputs=200, p=0.02, input_rate=500.0,
                            whi=3e-3, dmin=2, dmax=6, monitors=(0,))
    S = d.n_synapses
    syn = np.sort(np.random.default_rng(71).choice(S, size=S // 2, replace=False)).astype(np.int64)
    d.rules.append(Rule(RULE_WD, syn, builders.wd_params(lam=1e-2, alpha=2.02, w_max=4e-3)))
    gpu, ora = _run_both(d, 150, partitions=1, expect_fused=False)
    _compare(gpu, ora, d, "wd csr order")


@pytest.mark.parametrize("kind", [0, SYN_CONDUCTANCE])
def test_generator_inputs(kind):
    """Scheduled input spikes (three stimuli, random ids and times, an entry repeated
    within one step, entries after the run, an onset adjustment) on the per-phase path,
    with a stimulus switch: bit-exact against the oracle."""
    d = builders.random_net(seed=81, n_neurons=120, n_inputs=40, kind=kind, p=0.2, whi=4e-3, dmin=2,
                            dmax=6, monitors=(0, 1))
    if kind == SYN_CONDUCTANCE:
        d.weight = (d.weight * 400).astype(np.float32)
    rng = np.random.default_rng(81)
    stim = []
    for k in range(3):
        n = 400 + 50 * k
        ids = rng.integers(0, d.n_inputs, n).astype(np.int32)
        times = rng.uniform(0.0, 0.07, n).astype(np.float32)
        ids[:3] = ids[3]
        times[:3] = times[3] + np.float32(1e-6) * np.arange(3, dtype=np.float32)
        stim.append((ids, times))
    d.stimuli = stim
    d.onset = 5e-4
    print(f"[parity] generator inputs kind={kind}", flush=True)
    gpu = abi.Network(d)
    ora = oracle_py.OracleNetwork(d, precision=32)
    assert gpu.info().fused == 0 and "generator" in gpu.fused_status()
    ng = 150
    gpu.run(0.0, ng)
    ora.run(0.0, ng)
    _compare(gpu, ora, d, "generator stimulus 0")
    assert gpu.counters().spikes_inputs > 0
    assert abi.check(gpu.lib.spk_select_stimulus(gpu.h, 2)) is None
    assert ora.lib.or_select_stimulus(ora.h, 2) == 0
    t = float(d.group_time(0.0, ng))
    gpu.run(t, ng)
    ora.run(t, ng)
    _compare(gpu, ora, d, "generator stimulus 2")


@pytest.mark.parametrize("path", ["fused", "launches", "generator"])
def test_rate_monitor_counts(path):
    """RateActivityMonitor: per-unit spike counts kept by the kernels that make the
    spikes equal the bin counts of the spike monitors' rasters (which equal the oracle's),
    for the neurons and the inputs, on the persistent and per-phase paths and with
    generator inputs; a reset restarts one monitor's counts only."""
    d = builders.random_net(seed=91, input_rate=400.0, whi=4e-3, dmin=2, dmax=7, monitors=(0, 1))
    if path == "generator":
        rng = np.random.default_rng(91)
        d.stimuli = [(rng.integers(0, d.n_inputs, 300).astype(np.int32),
                      rng.uniform(0.0, 0.05, 300).astype(np.float32))]
    d.rate_monitors = [0, 1, 0]
    print(f"[parity] rate monitors path={path}", flush=True)
    gpu = abi.Network(d)
    ora = oracle_py.OracleNetwork(d, precision=32)
    launches = path != "fused"
    assert gpu.info().fused == (0 if path == "generator" else 1)
    gpu.run(0.0, 80, force_launches=launches)
    ora.run(0.0, 80)
    oi, _ = ora.take_spikes(0)
    oin, _ = ora.take_spikes(1)
    gi, _ = gpu.take_spikes(0)
    gin, _ = gpu.take_spikes(1)
    assert np.array_equal(np.sort(gi), np.sort(oi)) and np.array_equal(np.sort(gin), np.sort(oin))
    assert len(oi) > 0 and len(oin) > 0
    cn = np.bincount(oi, minlength=d.n_neurons)
    assert np.array_equal(gpu.rate_counts(0), cn)
    assert np.array_equal(gpu.rate_counts(1), np.bincount(oin, minlength=d.n_inputs))
    gpu.rate_reset(2)
    t = float(d.group_time(0.0, 80))
    gpu.run(t, 60, force_launches=launches)
    ora.run(t, 60)
    oi2, _ = ora.take_spikes(0)
    c2 = np.bincount(oi2, minlength=d.n_neurons)
    assert np.array_equal(gpu.rate_counts(0), cn + c2)
    assert np.array_equal(gpu.rate_counts(2), c2)


@pytest.mark.parametrize("labels", [1, 3])
@pytest.mark.parametrize("fused", [True, False])
def test_current_synapses(labels, fused):
    """Exponential current synapses (CurrentSpikingSynapses with the reference kernel's
    defects fixed): per label the current trace decays and takes the slot's input, and
    the neuron receives R' times their sum.  Persistent and per-phase paths."""
    d = builders.random_net(seed=95 + labels, kind=2, n_labels=labels, input_rate=500.0, whi=6e-10,
                            scale=1.0, dmin=2, dmax=6)
    gpu, ora = _run_both(d, 150, partitions=5, launches=not fused)
    assert gpu.info().synapse_kind == 2
    assert ora.counters()["spikes_neurons"] > 0
    _compare(gpu, ora, d, f"current L={labels}")


def test_time_only_advances_and_rules_list_synapses_once():
    d = _plastic_net(RULE_WD, 70)
    gpu = abi.Network(d)
    gpu.run(0.0, 20)
    with pytest.raises(abi.SpikeError, match="precedes"):
        gpu.run(0.0, 5)
    with pytest.raises(abi.SpikeError, match="precedes"):
        gpu.step_phases(float(d.group_time(0.0, 3)))
    gpu.reset_state()
    gpu.run(0.0, 5)
    gpu.close()
    syn = d.rules[0].synapses
    d.rules[0].synapses = np.concatenate([syn, syn[:1]])
    with pytest.raises(abi.SpikeError, match="twice"):
        abi.Network(d)


@pytest.mark.parametrize("fmt", [WEIGHTS_BF16, WEIGHTS_FP16])
def test_conductance_16bit_weights(fmt):
    d = builders.random_net(seed=17, kind=SYN_CONDUCTANCE, n_labels=2, whi=1.0,
                            scale=lambda rng, S: rng.uniform(0.5e-9, 2e-9, S), input_rate=500.0,
                            dmin=2, dmax=5)
    d.weight_format = fmt
    gpu, ora = _run_both(d, 250)
    _compare(gpu, ora, d, f"conductance, weights format {fmt}")


@pytest.mark.parametrize("fmt", [WEIGHTS_FP32, WEIGHTS_FP16])
def test_weight_normalization_after_stdp(fmt):
    """WeightNormSTDPPlasticity after a WD rule on the same synapses (per-phase path):
    every array bit-exact against the oracle, FP32 and binary16 weights."""
    from spike_amd.netdesc import RULE_WEIGHTNORM
    d = _plastic_net(RULE_WD, 47)
    d.weight_format = fmt
    syn = d.rules[0].synapses
    w0 = d.weight[syn].astype(np.float64)
    sum_sq = np.zeros(d.n_neurons, np.float64)
    np.add.at(sum_sq, d.post[syn], w0 ** 2)
    in_set = np.zeros(d.n_neurons, bool)
    in_set[d.post[syn]] = True
    d.rules.append(Rule(RULE_WEIGHTNORM, syn, dict(sum_sq=sum_sq.astype(np.float32), in_set=in_set)))
    gpu, ora = _run_both(d, 200, expect_fused=False)
    assert "normalization" in gpu.fused_status()
    _compare(gpu, ora, d, f"weight normalization fmt {fmt}")
    assert not np.array_equal(gpu.weights()[syn], d.weight[syn])


def test_custom_nearest_spike():
    d = _plastic_net(RULE_CUSTOM, 44)
    d.rules[0].params["nearest_spike_only"] = 1
    gpu, ora = _run_both(d, 200)
    _compare(gpu, ora, d, "custom nearest")
    assert not np.array_equal(gpu.weights(), d.weight), "plasticity changed no weight"


def test_wd_nearest_spike():
    d = _plastic_net(RULE_WD, 40)
    d.rules[0].params["nearest_spike_only"] = 1
    gpu, ora = _run_both(d, 200)
    _compare(gpu, ora, d, "wd nearest")


def test_coba_vogels_small():
    d = builders.coba_vogels(n_exc=320, n_inh=80, p=0.1, seed=5)
    gpu, ora = _run_both(d, 100)
    _compare(gpu, ora, d, "coba")


def test_weights_roundtrip_and_errors():
    d = builders.random_net(seed=9)
    gpu = abi.Network(d)
    w = gpu.weights()
    assert np.array_equal(w, d.weight)
    w2 = (w * 2).astype(np.float32)
    gpu.set_weights(w2)
    assert np.array_equal(gpu.weights(), w2)
    with pytest.raises(abi.SpikeError):
        gpu.set_grouping(100)
    with pytest.raises(abi.SpikeError):
        abi.check(gpu.lib.spk_update_plasticity(gpu.h, 5, 0.0))


@pytest.mark.parametrize("launches", [False, True])
def test_step_group_async(launches):
    """Groups enqueued one at a time (the backend's per-group call) give the oracle's run;
    monitor records stay in the device stores until drained."""
    d = _plastic_net(RULE_WD, 50)
    d.monitors = [0, 1]
    ng = 120
    print(f"[parity] step_group {d.name} groups={ng} launches={launches}", flush=True)
    gpu = abi.Network(d, partitions=5)
    ora = oracle_py.OracleNetwork(d, precision=32)
    for s in range(ng):
        gpu.step_group(d.group_time(0.0, s), plasticity=True, force_launches=launches)
    gpu.synchronize()
    ora.run(0.0, ng)
    c = gpu.counters()
    oc = ora.counters()
    assert c.spikes_neurons == oc["spikes_neurons"] and c.synaptic_events == oc["synaptic_events"]
    _compare(gpu, ora, d, f"step_group launches={launches}")


def test_run_then_steps_then_run():
    """spk_run, single-group steps and spk_run again continue one trajectory."""
    d = builders.random_net(seed=21, input_rate=400.0, whi=4e-3, dmin=2, dmax=7, ext_rate=2000.0,
                            ext_w=1e-3)
    print(f"[parity] run/steps/run {d.name}", flush=True)
    gpu = abi.Network(d, partitions=4)
    ora = oracle_py.OracleNetwork(d, precision=32)
    gpu.run(0.0, 40)
    for s in range(40, 70):
        gpu.step_group(d.group_time(0.0, s))
    gpu.run(d.group_time(0.0, 70), 50)
    ora.run(0.0, 70)
    ora.run(d.group_time(0.0, 70), 50)
    _compare(gpu, ora, d, "run/steps/run")
