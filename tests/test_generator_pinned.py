"""Generator inputs (GeneratorInputSpikingNeurons.cu:71-128) in the oracle, pinned by hand:
a scheduled spike at time T fires at the step n with |n*dt - T| < dt/2 (shifted by the
onset adjustment), sets the input's history bit and last spike time, is recorded once by
the input monitor, and reaches its target delay steps later with the synapse's weight;
two entries for one step fire once; select_stimulus replays another stimulus."""
import numpy as np

import oracle_py
from spike_amd.netdesc import F32, NetDesc, SYN_VOLTAGE, lif_constants


def _net(stimuli, onset=0.0, delay=3):
    d = NetDesc(n_neurons=2, n_inputs=4, kind=SYN_VOLTAGE)
    d.v0 = np.full(2, F32(-70e-3), F32)
    d.vreset = np.full(2, F32(-70e-3), F32)
    d.threshold = np.full(2, F32(-50e-3), F32)
    d.decay_c, d.resist = lif_constants(2, 200e-12, 10e-9, d.dt)
    d.pre = np.array([-1, -2], np.int32)          # inputs 0 and 1
    d.post = np.array([0, 1], np.int32)
    d.weight = np.array([1e-4, 2e-4], F32)
    d.delay = np.array([delay, delay], np.int32)
    d.grouping = 1
    d.stimuli = [(np.asarray(i, np.int32), np.asarray(t, F32)) for i, t in stimuli]
    d.onset = onset
    d.monitors = [1]
    return d


def _raster(o):
    ids, t = o.take_spikes(0)
    steps = np.rint(t.astype(np.float64) / 1e-4).astype(int)
    order = np.lexsort((ids, steps))
    return list(zip(steps[order].tolist(), ids[order].tolist()))


def test_scheduled_spikes_fire_at_their_steps_once():
    d = _net([([0, 1, 1, 2, 3], [0.0012, 0.0020, 0.00201, 0.0031, 0.05])])
    o = oracle_py.OracleNetwork(d, precision=32)
    o.run(0.0, 40)
    # input 1 has two entries within one step: one spike; input 3 is scheduled after the run
    assert _raster(o) == [(12, 0), (20, 1), (31, 2)]
    last = o.last_spike(1)
    assert last[0] == F32(12) * F32(1e-4)      # group time of step 12 (G = 1, sub-step 0)
    assert last[1] == F32(20) * F32(1e-4)
    assert last[3] == F32(-1000.0)
    assert o.counters()["spikes_inputs"] == 3


def test_delivery_after_the_delay():
    d = _net([([0], [0.0012])], delay=3)
    o = oracle_py.OracleNetwork(d, precision=32)
    o.run(0.0, 14)                      # steps 0..13: the spike at 12 is in the ring for step 15
    frac = o.info()["frac"]
    ring = o.ring().reshape(o.info()["B"], 2)
    slot = 15 % o.info()["B"]
    assert ring[slot, 0] == np.rint(np.float64(F32(1e-4)) * 2.0 ** frac)
    assert ring[:, 1].sum() == 0


def test_onset_and_stimulus_switch():
    d = _net([([0], [0.0012]), ([2], [0.0030])], onset=0.001)
    o = oracle_py.OracleNetwork(d, precision=32)
    o.run(0.0, 30)
    assert _raster(o) == [(22, 0)]      # t - onset = 0.0012 at step 22
    assert o.lib.or_select_stimulus(o.h, 1) == 0
    o.run(float(F32(30e-4)), 20)
    assert _raster(o) == [(40, 2)]      # t - onset = 0.0030 at step 40
