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


def test_voltage_injection_step_by_step():
    """Delta synapses (InlineDeviceFunctions.hpp:86-111 with LIFSpikingNeurons.cu:168-175):
    two generator inputs with delays 2 and 5 into one neuron; every step the membrane
    takes c*v0 + (1-c)*V_old + c*I_bg + (sum of the step's ring slot), recomputed in
    float32, and the slot is emptied once read -- a slot left full would deliver again one
    ring turn (B steps) later.  One arrival crosses threshold; the arrivals inside the
    following 2 ms are consumed and lost, and V stays at V_reset."""
    d = NetDesc(n_neurons=1, n_inputs=2, kind=SYN_VOLTAGE)
    d.v0 = np.full(1, F32(-70e-3), F32)
    d.vreset = np.full(1, F32(-75e-3), F32)
    d.threshold = np.full(1, F32(-60e-3), F32)
    d.decay_c, d.resist = lif_constants(1, 200e-12, 10e-9, d.dt)
    d.refractory = float(F32(2e-3))
    d.pre = np.array([-1, -2], np.int32)
    d.post = np.array([0, 0], np.int32)
    w = [F32(2.0 ** -9), F32(2.0 ** -6)]            # ~2 mV and ~16 mV, exact in the ring
    d.weight = np.array(w, F32)
    d.delay = np.array([2, 5], np.int32)
    fire = [[3, 4, 30, 31, 60], [20, 24]]
    ids = np.concatenate([np.full(len(f), l, np.int32) for l, f in enumerate(fire)])
    times = np.concatenate([np.array([F32(k) * F32(1e-4) for k in f], F32) for f in fire])
    order = np.argsort(times, kind="stable")
    d.stimuli = [(ids[order], times[order])]
    d.grouping = 1
    d.monitors = [0]
    o = oracle_py.OracleNetwork(d, precision=32)
    B = o.info()["B"]
    arrivals = {}
    for l, f in enumerate(fire):
        for k in f:
            arrivals[k + int(d.delay[l])] = arrivals.get(k + int(d.delay[l]), F32(0.0)) + w[l]
    c, v0, vr, thr = d.decay_c[0], d.v0[0], d.vreset[0], d.threshold[0]
    one = F32(1.0)
    V, last, spikes = v0, F32(-1000.0), []
    for step in range(60 + 3 * B):
        t = F32(step) * F32(1e-4)
        o.run(float(t), 1)
        dv = arrivals.get(step, F32(0.0))
        if F32(F32(t - last) - F32(d.refractory)) > F32(F32(0.5) * F32(1e-4)):
            V = F32(F32(F32(c * v0) + F32(F32(one - c) * V)) + F32(c * F32(0.0))) + dv
            if V >= thr:
                spikes.append(step)
                last, V = t, vr
        assert o.membrane()[0] == np.float64(V), step
    got = np.rint(o.take_spikes(0)[1].astype(np.float64) / 1e-4).astype(int).tolist()
    assert got == spikes == [25]
    # the arrival at step 29 (input 1 fired at 24, delay 5) fell inside the refractory
    # window of the spike at 25 and was lost; nothing arrived again B steps later
    assert 29 in arrivals and 25 + 20 > 29


def _conductance_net(w, delays, fire_steps, taus=(5e-3, 10e-3), rev=(0.0, -80e-3), n_steps=None):
    """One LIF neuron (threshold out of reach) fed by two generator inputs through two
    conductance labels: input l fires at the steps in fire_steps[l] and reaches label l
    after delays[l] steps."""
    from spike_amd.netdesc import conductance_decay
    d = NetDesc(n_neurons=1, n_inputs=2, kind=SYN_CONDUCTANCE, n_labels=2)
    d.v0 = np.full(1, F32(-70e-3), F32)
    d.vreset = np.full(1, F32(-70e-3), F32)
    d.threshold = np.full(1, F32(1.0), F32)
    d.decay_c, d.resist = lif_constants(1, 200e-12, 10e-9, d.dt)
    d.decay_g = np.array([conductance_decay(d.dt, t) for t in taus], F32)
    d.reversal = np.array(rev, F32)
    d.pre = np.array([-1, -2], np.int32)
    d.post = np.array([0, 0], np.int32)
    d.label = np.array([0, 1], np.int32)
    d.weight = np.array(w, F32)
    d.delay = np.array(delays, np.int32)
    ids = np.concatenate([np.full(len(s), l, np.int32) for l, s in enumerate(fire_steps)])
    times = np.concatenate([np.array([F32(k) * F32(1e-4) for k in s], F32) for s in fire_steps])
    order = np.argsort(times, kind="stable")
    d.stimuli = [(ids[order], times[order])]
    d.grouping = 1
    d.monitors = [0]
    return d


def test_conductance_injection_step_by_step():
    """Two conductance labels with opposite reversal potentials (0 mV and -80 mV) on one
    neuron, recomputed step by step in float32 (InlineDeviceFunctions.hpp:22-51 with
    LIFSpikingNeurons.cu:168-169): per label g = d*g + input, then
    I += g*(E - V_old), and the membrane takes c*v0 + (1-c)*V_old + c*I_bg + R'*I.  The
    conductances also follow the closed form g(n) = g(n0)*d^(n-n0) between arrivals, and
    each label pushes V towards its own reversal potential."""
    w = [F32(2.0 ** -30), F32(2.0 ** -28)]          # exact in the fixed-point ring
    delays = [1, 2]
    fire = [[5, 6, 20, 40], [8, 20, 21]]
    arrive = [{s + delays[l] for s in fire[l]} for l in range(2)]
    d = _conductance_net(w, delays, fire)
    o = oracle_py.OracleNetwork(d, precision=32)
    dec, rev = d.decay_g, d.reversal
    c, rp, v0 = d.decay_c[0], d.resist[0], d.v0[0]
    one = F32(1.0)
    g = [F32(0.0), F32(0.0)]
    V = v0
    last_arrival = [0, 0]
    g_at = [F32(0.0), F32(0.0)]
    for step in range(80):
        o.run(float(F32(step) * F32(1e-4)), 1)
        total = F32(0.0)
        for l in range(2):
            inp = w[l] if step in arrive[l] else F32(0.0)
            g[l] = F32(F32(dec[l] * g[l]) + inp)
            total = F32(total + F32(g[l] * F32(rev[l] - V)))
            if step in arrive[l]:
                last_arrival[l], g_at[l] = step, g[l]
        V_new = F32(F32(F32(c * v0) + F32(F32(one - c) * V)) + F32(c * F32(0.0))) + F32(total * rp)
        assert np.array_equal(o.conductances(), np.array(g, np.float64)), step
        assert o.membrane()[0] == np.float64(V_new), step
        for l in range(2):
            if step > last_arrival[l] and g_at[l] > 0:
                k = step - last_arrival[l]
                ref = float(g_at[l]) * float(dec[l]) ** k
                assert abs(float(g[l]) - ref) <= 1e-6 * ref, (step, l)
        V = V_new
    # each label pushes V towards its own reversal potential: against the same network
    # with that label's weight set to zero, the excitatory arrival at step 6 (E = 0)
    # raises V, the inhibitory one at step 10 (E = -80 mV) lowers it
    def trace(wv):
        net = oracle_py.OracleNetwork(_conductance_net(wv, delays, fire), precision=64)
        vs = []
        for step in range(12):
            net.run(float(F32(step) * F32(1e-4)), 1)
            vs.append(net.membrane()[0])
        return np.array(vs)
    both, no_inh, no_exc = trace(w), trace([w[0], 0.0]), trace([0.0, w[1]])
    assert both[6] == no_inh[6] and no_inh[6] > no_exc[6]
    assert both[10] < no_inh[10] and no_exc[10] < no_exc[9]


def test_conductance_at_rest_reversal_is_silent():
    """A conductance whose reversal potential equals the resting potential injects
    nothing while V sits there: with both labels at E = v0 the FP64 membrane equals the
    run without any input, step for step (a dropped (E - V) factor would not)."""
    fire = [[2, 3, 4, 5, 9], [3, 7]]
    base = oracle_py.OracleNetwork(_conductance_net([0.0, 0.0], [1, 1], fire, rev=(-70e-3, -70e-3)),
                                   precision=64)
    shunt = oracle_py.OracleNetwork(_conductance_net([F32(2.0 ** -28), F32(2.0 ** -27)], [1, 1], fire,
                                                     rev=(-70e-3, -70e-3)), precision=64)
    for step in range(30):
        base.run(float(F32(step) * F32(1e-4)), 1)
        shunt.run(float(F32(step) * F32(1e-4)), 1)
        assert shunt.membrane()[0] == base.membrane()[0], step
    assert shunt.conductances().max() > 0


def test_conductance_steady_state_is_the_weighted_reversal_mean():
    """Constant input on both labels every step: g_l settles at w_l / (1 - d_l) and V at
    the conductance-weighted mean of the reversal potentials,
    V* = (c*v0 + R'*sum_l g_l E_l) / (c + R'*sum_l g_l)  (the fixed point of
    V' = c*v0 + (1-c)*V + R'*sum_l g_l (E_l - V)), in FP64."""
    steps = 4000
    fire = [list(range(steps)), list(range(steps))]
    w = [F32(2.0 ** -32), F32(2.0 ** -33)]
    d = _conductance_net(w, [1, 1], fire)
    o = oracle_py.OracleNetwork(d, precision=64)
    o.run(0.0, steps)
    dec = d.decay_g.astype(np.float64)
    gs = np.array([float(w[0]), float(w[1])]) / (1.0 - dec)
    assert np.allclose(o.conductances(), gs, rtol=1e-12, atol=0)
    c, rp, v0 = float(d.decay_c[0]), float(d.resist[0]), float(d.v0[0])
    E = d.reversal.astype(np.float64)
    vstar = (c * v0 + rp * float(np.dot(gs, E))) / (c + rp * float(gs.sum()))
    assert abs(o.membrane()[0] - vstar) < 1e-12
    assert -80e-3 < vstar < -45e-3 and abs(vstar - v0) > 5e-3


def _isolated_driven(n, lam=2.0, w=1e-3, seed=11):
    """n unconnected LIF neurons (threshold out of reach) with the per-neuron external
    Poisson drive of mean lam events per step, each event worth w volts."""
    d = NetDesc(n_neurons=n, n_inputs=0, kind=SYN_VOLTAGE, seed=seed)
    d.v0 = np.zeros(n, F32)
    d.vreset = np.zeros(n, F32)
    d.threshold = np.full(n, F32(1e3), F32)
    d.decay_c, d.resist = lif_constants(n, 200e-12, 10e-9, d.dt)
    d.ext_rate = float(lam / d.dt)
    d.ext_weight = float(F32(w))
    d.grouping = 1
    return d


def _drive_counts(o, d, prev):
    """Per-neuron event count of the last step, recovered from the FP64 membrane
    V' = c*v0 + (1-c)*V + k*w (v0 = 0)."""
    c = float(d.decay_c[0])
    v = o.membrane()
    k = (v - (1.0 - c) * prev) / d.ext_weight
    ki = np.rint(k)
    assert np.max(np.abs(k - ki)) < 1e-6
    return ki.astype(np.int64), v


def test_external_drive_count_is_poisson_inverse_cdf_of_philox():
    """The external drive of neuron i at step n is k = F^-1((u+1)/2^32) for the Poisson
    (lambda) CDF F, u = draw n of (i, stream 2) -- output n mod 4 of the Philox call at
    counter (n/4, 0, i, 2), key = seed (KAT-checked above) -- and k events of weight w
    join the step's input.  Checked per neuron against scipy's Poisson quantile over two
    steps of 4000 neurons."""
    from scipy.stats import poisson
    lam, seed, n = 2.0, 11, 4000
    d = _isolated_driven(n, lam=lam, seed=seed)
    o = oracle_py.OracleNetwork(d, precision=64)
    prev = np.zeros(n)
    for step in (0, 1):
        o.run(float(F32(step) * F32(1e-4)), 1)
        k, prev = _drive_counts(o, d, prev)
        u = np.array([oracle_py.philox([step // 4, 0, i, 2], [seed & 0xFFFFFFFF, seed >> 32])[step % 4]
                      for i in range(n)], np.float64)
        expect = poisson.ppf((u + 1.0) / 2.0 ** 32, lam).astype(np.int64)
        assert np.array_equal(k, expect), step
    assert o.counters()["external_events"] > 0


def test_external_drive_count_histogram_is_poisson():
    """Over 10^5 (neuron, step) samples the per-step event counts follow the Poisson(2)
    pmf: every bin with expectation above 20 lies within 5 sigma, and the sample mean
    and variance are both lambda within sampling error."""
    lam, n, steps = 2.0, 20000, 5
    d = _isolated_driven(n, lam=lam, seed=5)
    o = oracle_py.OracleNetwork(d, precision=64)
    prev = np.zeros(n)
    ks = []
    for step in range(steps):
        o.run(float(F32(step) * F32(1e-4)), 1)
        k, prev = _drive_counts(o, d, prev)
        ks.append(k)
    ks = np.concatenate(ks)
    m = len(ks)
    hist = np.bincount(ks, minlength=16)
    for j in range(16):
        e = m * math.exp(-lam) * lam ** j / math.factorial(j)
        if e > 20:
            assert abs(hist[j] - e) < 5 * math.sqrt(e), (j, hist[j], e)
    assert abs(ks.mean() - lam) < 5 * math.sqrt(lam / m)
    assert abs(ks.var() - lam) < 5 * lam * math.sqrt(2.0 / m) + 0.02
    assert o.counters()["external_events"] == int(ks.sum())


def test_brunel_connectivity_counts_and_rows():
    """The Philox fixed in-degree rule (the sweep's generated networks): every neuron has
    exactly C_E sources in [0, N_E) followed by C_I in [N_E, N); slot j of neuron i is
    floor(u*N_E/2^32) (j < C_E) or N_E + floor(u*N_I/2^32), u = draw j of (i, stream 3),
    recomputed for a few rows from the KAT-checked block function; and the sources of
    each type are uniform (chi-square over all neurons)."""
    seed, ne, ni, ce, ci = 2024, 1600, 400, 160, 40
    pre = oracle_py.brunel_connectivity(seed, ne, ni, ce, ci)
    n = ne + ni
    assert pre.shape == (n, ce + ci)
    assert np.all((pre[:, :ce] >= 0) & (pre[:, :ce] < ne))
    assert np.all((pre[:, ce:] >= ne) & (pre[:, ce:] < n))
    key = [seed & 0xFFFFFFFF, seed >> 32]
    for i in (0, 1, 777, n - 1):
        row = []
        for j in range(ce + ci):
            u = int(oracle_py.philox([j // 4, 0, i, 3], key)[j % 4])
            row.append((u * ne) >> 32 if j < ce else ne + ((u * ni) >> 32))
        assert pre[i].tolist() == row, i
    from scipy.stats import chisquare
    he = np.bincount(pre[:, :ce].ravel(), minlength=ne)
    hi = np.bincount(pre[:, ce:].ravel() - ne, minlength=ni)
    assert chisquare(he).pvalue > 1e-4 and chisquare(hi).pvalue > 1e-4
    # rows are independent draws: different neurons do not share their source lists
    assert len({tuple(r) for r in pre[:50]}) == 50
    # another seed gives another network
    assert not np.array_equal(pre, oracle_py.brunel_connectivity(seed + 1, ne, ni, ce, ci))


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
