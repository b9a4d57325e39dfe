"""FP64 oracle statistics of the COBA + Vogels iSTDP network (3200 E + 800 I, p = 0.02,
2 s, or the configuration's 10 s) for connectivity seeds 42..49: spike count, mean ISI CV,
the final I->E weights' mean and 20-bin histogram (edges from their range), and the
histogram of the same weights rounded to bfloat16.  Writes coba_vogels_fp64_<seconds>s.json
beside this script.

    python tests/golden/make_coba_fp64.py [threads] [seconds]
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")]

import numpy as np  # noqa: E402

import oracle_py  # noqa: E402
from chaos_stats import histogram, isi_cv, round_bf16  # noqa: E402
from spike_amd import builders  # noqa: E402

SEEDS = list(range(42, 50))
BINS = 20


def fp64_entry(seed, threads, seconds):
    d = builders.coba_vogels(seed=seed, monitors=(0,))
    o = oracle_py.OracleNetwork(d, precision=64, threads=threads)
    o.run(0.0, d.groups_for(seconds))
    ids, times = o.take_spikes(0)
    w = o.weights()[d.rules[0].synapses]
    o.close()
    edges = np.linspace(w.min(), w.max(), BINS + 1)
    return dict(spikes=int(len(ids)), cv=isi_cv(ids, times, d.n_neurons), w_mean=float(w.mean()),
                w_std=float(w.std()), edges=edges.tolist(), hist=histogram(w, edges).tolist(),
                hist_bf16=histogram(round_bf16(w.astype(np.float32)), edges).tolist())


def main():
    threads = int(sys.argv[1]) if len(sys.argv) > 1 else 8
    seconds = float(sys.argv[2]) if len(sys.argv) > 2 else 2.0
    out = dict(seconds=seconds, network="builders.coba_vogels(seed)", seeds={})
    for seed in SEEDS:
        out["seeds"][str(seed)] = e = fp64_entry(seed, threads, seconds)
        print(seed, e["spikes"], round(e["cv"], 4), flush=True)
    with open(os.path.join(HERE, "coba_vogels_fp64_%gs.json" % seconds), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
