"""Expected output of Examples/Brunel10K.cpp run unchanged on the HIP backend for 0.2 s
(`Brunel10K_hip --simtime 0.2`): the FP32 oracle's spike count of the same network.

The network is rebuilt with the oracle's restatement of the front-end's construction
(oracle/construction.py, pinned draw for draw against fixtures the reference front-end
wrote): srand(42) from `VoltageSpikingSynapses(42)`, then six `connect_with_sparsity`
projections in the example's order (input->E, input->I, E->I, E->E, I->E, I->I; fixed
in-degree, sparseness 0.1, sources `rand() % n_pre`, Examples/UtilityFunctions.hpp:51-82),
weights 0.1 mV (inhibitory -5 x 0.1 mV), every delay 1.5 ms; 8000 E + 2000 I LIF (200 pF,
10 nS, V0 = Vreset = 0, threshold 20 mV, 2 ms refractory); 10^4 Poisson inputs at 20 Hz
with the backend's default Philox seed (1).  Writes brunel10k_example_0p2s.json.

    python tests/golden/make_brunel_example_fixture.py [threads]
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle")]

import numpy as np  # noqa: E402

from construction import Builder  # noqa: E402
from spike_amd import netdesc  # noqa: E402

F32 = np.float32
SIMTIME = 0.2


def example_network():
    dt = F32(0.0001)
    b = Builder(42, timestep=0.0001)
    inp = b.add_inputs(10000)
    exc = b.add_neurons(8000)
    inh = b.add_neurons(2000)
    delay = F32(F32(1.5) * F32(np.power(10.0, -3.0)))
    w = F32(F32(0.1) * F32(np.power(10.0, -3.0)))
    wi = F32(F32(-5.0) * w)
    sp = 0.1
    for pre, post, n_pre, n_post, wv in ((inp, exc, 10000, 8000, w), (inp, inh, 10000, 2000, w),
                                         (exc, inh, 8000, 2000, w), (exc, exc, 8000, 8000, w),
                                         (inh, exc, 2000, 8000, wi), (inh, inh, 2000, 2000, wi)):
        b.connect_with_sparsity(pre, post, n_pre, n_post, sp, weight_range=(wv, wv), scale=1.0,
                                delay_range=(delay, delay))
    f = b.finalise()
    N = 10000
    # the example's `200.0f*pow(10.0, -12)` etc.: float times double, rounded to float
    cm, g0 = F32(200.0 * 10.0 ** -12), F32(10.0 * 10.0 ** -9)
    thr, tref = F32(20.0 * 10.0 ** -3), F32(2.0 * 10.0 ** -3)
    c, rp = netdesc.lif_constants(N, cm, g0, float(dt))
    d = netdesc.NetDesc(
        n_neurons=N, n_inputs=10000, kind=netdesc.SYN_VOLTAGE, dt=float(dt),
        v0=np.zeros(N, F32), vreset=np.zeros(N, F32), threshold=np.full(N, thr, F32),
        decay_c=c, resist=rp, refractory=float(tref), background=0.0,
        rates=np.full((1, 10000), F32(20.0), F32), seed=1,
        pre=f["pre"], post=f["post"], weight=f["w"], scale=f["scale"], delay=f["delay"], label=f["label"],
        monitors=[netdesc.POP_NEURONS], grouping=f["timestep_grouping"], name="brunel10k_example")
    return d


def main():
    threads = int(sys.argv[1]) if len(sys.argv) > 1 else 8
    import oracle_py
    t0 = time.time()
    d = example_network()
    t1 = time.time()
    o = oracle_py.OracleNetwork(d, precision=32, threads=threads)
    o.run(0.0, d.groups_for(SIMTIME))
    ids, times = o.take_spikes(0)
    o.close()
    out = dict(program="Examples/Brunel10K.cpp", args=["--simtime", str(SIMTIME)], simtime=SIMTIME,
               n_synapses=int(d.n_synapses), timestep_grouping=int(d.timestep_grouping()),
               groups=int(d.groups_for(SIMTIME)), spikes=int(len(ids)),
               spikes_per_neuron_sum_of_squares=int(np.sum(np.bincount(ids, minlength=d.n_neurons).astype(np.int64) ** 2)),
               build_s=round(t1 - t0, 1), oracle_s=round(time.time() - t1, 1))
    with open(os.path.join(HERE, "brunel10k_example_0p2s.json"), "w") as fh:
        json.dump(out, fh, indent=1)
    print(out)


if __name__ == "__main__":
    main()
