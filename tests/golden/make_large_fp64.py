"""FP64 oracle statistics of the large plastic network's configuration at 1/25 of its size
(2e4 LIF neurons, 80/20, 1000 Philox fixed out-degree synapses each; 1000 Poisson inputs
at 10 Hz with 1000 targets each, so every neuron keeps the full-size in-degrees: ~800 E,
~200 I, ~50 input synapses; nearest-spike weight-dependent STDP on E->E, 2 s), for seeds
1..4: spike count, mean ISI CV, the final E->E weights' mean, std and 20-bin histogram
over [0, 1].  Writes large_plastic_fp64_2s.json beside this script.

    python tests/golden/make_large_fp64.py [threads] [procs]

The full-size network (5.25e8 synapses) would take the oracle about six hours per 2 s;
the in-degrees, drive and rule, which set the dynamics, are those of the full size.
"""
import json
import multiprocessing as mp
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")]

import numpy as np  # noqa: E402

SEEDS = [1, 2, 3, 4]
N = 20000
SECONDS = 2.0
BINS = 20


def network(seed):
    from spike_amd import builders
    return builders.large_plastic(n=N, seed=seed, monitors=(0,))


def fp64_entry(job):
    seed, threads = job
    import oracle_py
    from chaos_stats import histogram, isi_cv
    d = network(seed)
    e = oracle_py.expand_outdegree(d, threads=threads)
    t0 = time.time()
    o = oracle_py.OracleNetwork(e, precision=64, threads=threads)
    o.run(0.0, d.groups_for(SECONDS))
    ids, times = o.take_spikes(0)
    w = o.weights()[e.rules[0].synapses]
    o.close()
    edges = np.linspace(0.0, 1.0, BINS + 1)
    return seed, dict(spikes=int(len(ids)), rate_hz=float(len(ids) / N / SECONDS), cv=isi_cv(ids, times, N),
                      w_mean=float(w.mean()), w_std=float(w.std()), edges=edges.tolist(),
                      hist=histogram(w, edges).tolist(), wall_s=round(time.time() - t0, 1))


def main():
    threads = int(sys.argv[1]) if len(sys.argv) > 1 else 3
    procs = int(sys.argv[2]) if len(sys.argv) > 2 else 2
    out = dict(seconds=SECONDS, network=f"builders.large_plastic(n={N}, seed)", n_neurons=N, seeds={})
    with mp.get_context("spawn").Pool(procs) as pool:
        for seed, e in pool.imap_unordered(fp64_entry, [(s, threads) for s in SEEDS]):
            out["seeds"][str(seed)] = e
            print(seed, e["spikes"], round(e["cv"], 4), round(e["w_mean"], 6), e["wall_s"], flush=True)
    out["seeds"] = dict(sorted(out["seeds"].items(), key=lambda kv: int(kv[0])))
    with open(os.path.join(HERE, "large_plastic_fp64_2s.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
