"""Fixed out-degree networks generated on the device (spk_set_outdegree_synapses), and the
large plastic network at its configured size.

1. A generated network equals, bit for bit, the FP32 oracle running the same network built
   on the host from the oracle's restatement of the generation rule, and the same network
   prepared by the library from those host arrays (spk_set_synapses): rasters, membranes,
   last-spike times, history, the input ring, weights and traces, FP32 and FP16 storage.
2. The large plastic network at its configured size (5e5 LIF, 1000 out-synapses each,
   2.5e4 Poisson inputs with 1000 each: 5.25e8 synapses, FP16 weights, WD STDP on the
   3.2e8 E->E synapses): its first 20 steps equal the oracle's bit for bit (synapse indices
   reach 5.25e8, byte offsets of the synapse records 4.2e9); over 2 s the device-generated
   and the host-prepared networks agree bit for bit, every spike reaches exactly its
   source's out-degree, the monitors hold every counted spike, plastic weights stay in
   [0, w_max] and the other weights keep their initial values.
"""
import time

import numpy as np
import pytest

import oracle_py
from spike_amd import abi, builders

from test_gpu_parity import _compare  # noqa: E402

pytestmark = pytest.mark.gpu
F32 = np.float32


def _info_tuple(i):
    return (i.timestep_grouping, i.ring_slots, i.history_words, i.frac_bits, i.min_delay, i.max_delay,
            i.n_synapses)


def _oracle_info(o, S):
    return (o["G"], o["B"], o["W"], o["frac"], o["dmin"], o["dmax"], S)


@pytest.mark.parametrize("fmt", [0, 2])
def test_generated_equals_oracle_and_host_path(fmt):
    d = builders.large_plastic(n=5000, out_degree=100, monitors=(0, 1))
    d.weight_format = fmt
    steps = 1000
    print(f"[outdegree] N=5000 K=100 Nin=250 Kin=1000 fmt={fmt}: {steps} steps", flush=True)
    e = oracle_py.expand_outdegree(d, threads=16)
    gpu = abi.Network(d)
    host = abi.Network(e)
    ora = oracle_py.OracleNetwork(e, precision=32, threads=16)
    gi = gpu.info()
    assert gi.fused == 0 and gi.n_synapses == d.n_synapses
    assert _info_tuple(gi) == _info_tuple(host.info()) == _oracle_info(ora.info(), e.n_synapses)
    assert gpu.rule_size(0) == host.rule_size(0) == len(e.rules[0].synapses)
    for net in (gpu, host):
        net.run(0.0, steps)
    ora.run(0.0, steps)
    c, oc = gpu.counters(), ora.counters()
    assert (c.spikes_neurons, c.spikes_inputs, c.synaptic_events) == \
        (oc["spikes_neurons"], oc["spikes_inputs"], oc["synaptic_events"])
    assert c.spikes_neurons > 0
    hc = host.counters()
    assert (hc.spikes_neurons, hc.synaptic_events) == (c.spikes_neurons, c.synaptic_events)
    print(f"  {c.spikes_neurons} neuron spikes, {c.synaptic_events} deliveries", flush=True)
    _compare(gpu, ora, e, f"generated fmt={fmt}")
    # the host-prepared network still holds its spikes (the oracle comparison drained gpu's)
    for m in range(2):
        ids, _ = host.take_spikes(m)
        assert len(ids) == (c.spikes_neurons if m == 0 else c.spikes_inputs)
    assert np.array_equal(gpu.weights(), host.weights())
    assert np.array_equal(gpu.membrane(), host.membrane())
    for n in (gpu, host, ora):
        n.close()


def test_generated_network_rejects_listed_rules_and_bad_sizes():
    d = builders.large_plastic(n=1000, out_degree=10, input_out_degree=10)
    lib = abi.load()
    h = abi.P()
    abi.check(lib.spk_create(1000, 50, 0, 1, 1e-4, abi.C.byref(h)))
    try:
        p = abi.OutDegreeParams(seed=1, n_exc=800, out_degree=10, input_out_degree=10, delay_max=0)
        assert lib.spk_set_outdegree_synapses(h, abi.C.byref(p)) != 0
        assert "delay_max" in lib.spk_last_error().decode()
        p.delay_max = 20
        p.n_exc = 1001
        assert lib.spk_set_outdegree_synapses(h, abi.C.byref(p)) != 0
        p.n_exc = 800
        p.out_degree = 1 << 23   # 1000 * 2^23 >= 2^32 - 1 synapses
        assert lib.spk_set_outdegree_synapses(h, abi.C.byref(p)) != 0
        p.out_degree = 10
        assert lib.spk_set_outdegree_synapses(h, abi.C.byref(p)) == 0
        syn = np.arange(3, dtype=np.int64)
        w = abi.WDParams(w_max=1.0)
        rid = abi.C.c_int()
        assert lib.spk_add_wd_stdp(h, abi._ptr(syn, abi.C.c_int64), 3, abi.C.byref(w), abi.C.byref(rid)) != 0
        assert "generated" in lib.spk_last_error().decode()
    finally:
        lib.spk_destroy(h)
    assert d.n_synapses == 1000 * 10 + 50 * 10


def _full_size(fmt=2, monitors=(0, 1)):
    d = builders.large_plastic(monitors=monitors)
    d.weight_format = fmt
    return d


@pytest.mark.timeout(900)
def test_large_plastic_full_size_first_20_steps_match_oracle():
    d = _full_size()
    print(f"[outdegree] full size: N={d.n_neurons} Nin={d.n_inputs} S={d.n_synapses}: generate", flush=True)
    t0 = time.perf_counter()
    gpu = abi.Network(d)
    gen_s = time.perf_counter() - t0
    print(f"  generated and prepared on the device in {gen_s:.2f} s; oracle restatement", flush=True)
    assert gen_s < 10.0
    e = oracle_py.expand_outdegree(d, threads=16)
    print("  oracle network", flush=True)
    ora = oracle_py.OracleNetwork(e, precision=32, threads=16)
    assert _info_tuple(gpu.info()) == _oracle_info(ora.info(), e.n_synapses)
    assert gpu.rule_size(0) == len(e.rules[0].synapses)
    steps = 20
    gpu.run(0.0, steps)
    print(f"  oracle: {steps} steps", flush=True)
    ora.run(0.0, steps)   # one run: group times start + s*G*dt as in the device run
    c, oc = gpu.counters(), ora.counters()
    assert (c.spikes_neurons, c.spikes_inputs, c.synaptic_events) == \
        (oc["spikes_neurons"], oc["spikes_inputs"], oc["synaptic_events"])
    assert c.spikes_inputs > 0
    _compare(gpu, ora, e, "large plastic full size, 20 steps")
    gpu.close()
    ora.close()


@pytest.mark.timeout(900)
def test_large_plastic_full_size_2s_generated_equals_host_path_and_properties():
    d = _full_size()
    o = d.outdegree
    steps = 20000
    print(f"[outdegree] full size 2 s: generated network", flush=True)
    gpu = abi.Network(d)
    gpu.run(0.0, steps)
    c = gpu.counters()
    ids, _ = gpu.take_spikes(0)
    iids, _ = gpu.take_spikes(1)
    print(f"  {c.spikes_neurons} neuron spikes ({c.spikes_neurons / d.n_neurons / 2.0:.2f} Hz), "
          f"{c.spikes_inputs} input spikes, {c.synaptic_events} deliveries", flush=True)
    assert c.spikes_neurons > 0
    assert len(ids) == c.spikes_neurons and len(iids) == c.spikes_inputs
    assert c.synaptic_events == c.spikes_neurons * o.out_degree + c.spikes_inputs * o.input_out_degree
    # weights: plastic ones in [0, w_max]; the others as stored at the start
    pre, post, w0, _, _ = oracle_py.outdegree_connectivity(d.seed, d.n_neurons, d.n_inputs, o, threads=16)
    w = gpu.weights()
    ee = (pre >= 0) & (pre < o.n_exc) & (post < o.n_exc)
    del pre, post
    assert np.all(np.isfinite(w))
    assert w[ee].min() >= 0.0 and w[ee].max() <= o.ee_stdp["w_max"]
    w16 = w0.astype(np.float16).astype(np.float32)   # FP16 storage of the initial weights
    assert np.array_equal(w[~ee], w16[~ee])
    moved = np.count_nonzero(w[ee] != w16[ee])
    print(f"  {moved} of {np.count_nonzero(ee)} plastic weights moved; mean {w[ee].mean():.6f} "
          f"(initial {w16[ee].mean():.6f})", flush=True)
    assert moved > 0
    del w0, w16, ee
    print("  host-prepared network of the same synapses", flush=True)
    e = oracle_py.expand_outdegree(d, threads=16)
    host = abi.Network(e)
    del e
    host.run(0.0, steps)
    hc = host.counters()
    assert (hc.spikes_neurons, hc.spikes_inputs, hc.synaptic_events) == \
        (c.spikes_neurons, c.spikes_inputs, c.synaptic_events)
    assert np.array_equal(host.weights(), w)
    assert np.array_equal(host.membrane(), gpu.membrane())
    hids, _ = host.take_spikes(0)
    assert np.array_equal(hids, ids)
    ta, tb = gpu.traces(0)
    ua, ub = host.traces(0)
    assert np.array_equal(ta, ua) and np.array_equal(tb, ub)
    gpu.close()
    host.close()
