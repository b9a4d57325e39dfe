"""Drop-in path: Spike's unchanged C++ front-end running on the HIP backend.

oracle/_ref/frontend_net_hip (tests/frontend/frontend_net.cpp linked with the reference's
front-end library and spike_amd/backend, default backend "HIP") builds a network through
SpikingModel / neuron groups / synapse groups / plasticity rules / SpikingActivityMonitor,
runs SpikingModel::run and saves the monitors' rasters with save_spikes_as_binary.  The
same network (its finalised arrays, dumped by the program) is replayed on the FP32 oracle;
rasters and plastic weights must match bit for bit.
"""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from spike_amd import netdesc  # noqa: E402

PROG = os.path.join(ROOT, "oracle", "_ref", "frontend_net_hip")
F32 = np.float32

pytestmark = pytest.mark.skipif(not os.path.exists(PROG), reason="front-end programs not built (needs the Spike source tree)")


def run_program(tmp_path, case, simtime, plastic):
    out = str(tmp_path / case)
    args = [PROG, case, out, str(simtime)] + (["plastic"] if plastic else [])
    print("frontend_net_hip", " ".join(args[1:]), flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=180)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    meta = json.load(open(os.path.join(out, "meta.json")))

    def arr(name, dt):
        return np.fromfile(os.path.join(out, name + ".bin"), dtype=dt)

    return meta, arr, r.stdout


def desc_from_dump(meta, arr):
    dt = F32(meta["timestep"])
    tau = arr("lif_tau", F32)
    R = arr("lif_R", F32)
    c = (dt / tau).astype(F32)
    rp = (c * R).astype(F32)
    n_in = meta["n_inputs"]
    cond = meta["conductance"] == 1
    curr = meta.get("current", 0) == 1
    kinds = {"wd": netdesc.RULE_WD, "vogels": netdesc.RULE_VOGELS, "evans": netdesc.RULE_EVANS,
             "custom": netdesc.RULE_CUSTOM, "weightnorm": netdesc.RULE_WEIGHTNORM}
    rules = []
    for r, (k, p) in enumerate(zip(meta["rule_kinds"], meta["rule_params"])):
        if k == "weightnorm":
            p = dict(sum_sq=arr("wn_sumsq_%d" % r, F32), in_set=arr("wn_inset_%d" % r, np.uint8).astype(bool))
        rules.append(netdesc.Rule(kinds[k], arr("plastic_%d" % r, np.int32).astype(np.int64), p))
    d = netdesc.NetDesc(
        n_neurons=meta["n_neurons"], n_inputs=n_in,
        kind=netdesc.SYN_CONDUCTANCE if cond else (netdesc.SYN_CURRENT if curr else netdesc.SYN_VOLTAGE),
        n_labels=max(1, meta["num_syn_labels"]) if (cond or curr) else 1, dt=float(dt),
        v0=arr("lif_v0", F32), vreset=arr("lif_vreset", F32), threshold=arr("lif_thr", F32),
        decay_c=c, resist=rp, refractory=meta["refractory"], background=meta["background_current"],
        rates=arr("rates", F32).reshape(1, n_in), seed=1,
        decay_g=arr("decay_g", F32) if (cond or curr) else None,
        reversal=arr("reversal", F32) if (cond or curr) else None,
        pre=arr("syn_pre", np.int32), post=arr("syn_post", np.int32), weight=arr("syn_w", F32),
        scale=arr("syn_scale", F32), delay=arr("syn_delay", np.int32), label=arr("syn_label", np.int32),
        rules=rules, monitors=[netdesc.POP_NEURONS, netdesc.POP_INPUTS],
        grouping=meta["timestep_grouping"], name="frontend_" + meta["case"])
    if meta.get("generator"):
        counts = arr("gen_counts", np.int32)
        ids, times = arr("gen_ids", np.int32), arr("gen_times", F32)
        off = np.concatenate([[0], np.cumsum(counts)])
        d.stimuli = [(ids[off[k]:off[k + 1]], times[off[k]:off[k + 1]]) for k in range(len(counts))]
    return d


def oracle_run(d, simtime, plastic):
    import oracle_py
    o = oracle_py.OracleNetwork(d, precision=32, threads=4)
    o.run(0.0, d.groups_for(simtime), plasticity=plastic)
    if d.stimuli is not None:
        # the program selects stimulus 1 and runs again from the current time
        # (SpikingModel::run continues at start + simtime, SpikingModel.cpp:214-236)
        assert o.lib.or_select_stimulus(o.h, 1) == 0
        o.run(float(F32(0.0) + F32(simtime)), d.groups_for(simtime), plasticity=plastic)
    n_ids, n_t = o.take_spikes(0)
    i_ids, i_t = o.take_spikes(1)
    w = o.weights().astype(np.float32)   # the FP32 oracle's weights, returned widened
    o.close()
    return (n_ids, n_t), (i_ids, i_t), w


def sort_raster(ids, times):
    order = np.lexsort((ids, times))
    return ids[order], times[order]


CASES = [("brunel", 0.1, False), ("brunel", 0.1, True), ("coba", 0.1, True), ("plain", 0.05, False),
         ("generator", 0.1, False), ("current", 0.1, False)]


@pytest.mark.gpu
@pytest.mark.parametrize("case,simtime,plastic", CASES)
def test_frontend_run_matches_oracle(tmp_path, case, simtime, plastic):
    meta, arr, stdout = run_program(tmp_path, case, simtime, plastic)
    d = desc_from_dump(meta, arr)
    (on_ids, on_t), (oi_ids, oi_t), ow = oracle_run(d, simtime, plastic)
    hn_ids, hn_t = sort_raster(arr("neurons_SpikeIDs", np.int32), arr("neurons_SpikeTimes", F32))
    hi_ids, hi_t = sort_raster(arr("inputs_SpikeIDs", np.int32), arr("inputs_SpikeTimes", F32))
    on_ids, on_t = sort_raster(on_ids, on_t)
    oi_ids, oi_t = sort_raster(oi_ids, oi_t)
    assert len(on_ids) > 0, "network is silent"
    assert meta["spikes_neurons"] == len(hn_ids)
    np.testing.assert_array_equal(hi_ids, oi_ids)
    np.testing.assert_array_equal(hi_t.view(np.uint32), oi_t.view(np.uint32))
    np.testing.assert_array_equal(hn_ids, on_ids)
    np.testing.assert_array_equal(hn_t.view(np.uint32), on_t.view(np.uint32))
    # the RateActivityMonitor on the neurons counted every spike of the run
    np.testing.assert_array_equal(arr("rate_counts", np.int32), np.bincount(on_ids, minlength=meta["n_neurons"]))
    w = arr("weights_after", F32)
    np.testing.assert_array_equal(w.view(np.uint32), ow.view(np.uint32))
    # Spike's connectivity files, read with spike_amd.io: the same synapses (creation
    # order) with the device's weights; loading them back leaves the device weights as they were
    from spike_amd import io
    conn = io.read_connectivity(str(tmp_path / case), "conn_")
    key = lambda p, q, dl, ww: np.lexsort((ww.view(np.uint32), dl, q, p))
    a = key(conn["pre"], conn["post"], conn["delay"], conn["weight"])
    b = key(arr("syn_pre", np.int32), arr("syn_post", np.int32), arr("syn_delay", np.int32), w)
    assert np.array_equal(conn["pre"][a], arr("syn_pre", np.int32)[b])
    assert np.array_equal(conn["delay"][a], arr("syn_delay", np.int32)[b])
    assert np.array_equal(conn["weight"][a].view(np.uint32), w[b].view(np.uint32))
    np.testing.assert_array_equal(arr("weights_reloaded", F32).view(np.uint32), w.view(np.uint32))
    if plastic:
        assert not np.array_equal(w, arr("syn_w", F32)), "plasticity changed no weight"


@pytest.mark.gpu
def test_brunel_example_unchanged_runs():
    """Examples/Brunel10K.cpp, compiled unchanged against the HIP backend, runs 0.2 s of
    the 10,000-neuron network with its spike monitor and prints the number of spikes
    recorded: exactly the FP32 oracle's count for the same network, rebuilt by the oracle's
    restatement of the front-end's construction (tests/golden/make_brunel_example_fixture.py,
    which wrote tests/golden/brunel10k_example_0p2s.json)."""
    gold = json.load(open(os.path.join(ROOT, "tests", "golden", "brunel10k_example_0p2s.json")))
    prog = os.path.join(ROOT, "oracle", "_ref", "Brunel10K_hip")
    print("Brunel10K_hip --simtime 0.2", flush=True)
    r = subprocess.run([prog, "--simtime", "0.2"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    lines = [l for l in r.stdout.splitlines() if "Number of Spikes Recorded" in l]
    assert lines, r.stdout[-2000:]
    n = int(lines[-1].split(":")[1])
    assert n == gold["spikes"] > 0


@pytest.mark.gpu
def test_simple_example_unchanged_runs(tmp_path):
    """Examples/SimpleExample.cpp, compiled unchanged against the HIP backend: generator
    inputs -> conductance LIF, 2 x 50 s.  Its input monitor records exactly the five
    scheduled spikes of the first stimulus (the second stimulus's times, 5-10 s, lie
    before the second run's start at 50 s, so it never fires)."""
    prog = os.path.join(ROOT, "oracle", "_ref", "SimpleExample_hip")
    if not os.path.exists(prog):
        pytest.skip("SimpleExample_hip not built")
    print("SimpleExample_hip", flush=True)
    r = subprocess.run([prog], capture_output=True, text=True, timeout=600, cwd=str(tmp_path))
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    ids = np.loadtxt(tmp_path / "SpikeIDs.txt", dtype=np.int64, ndmin=1)
    times = np.loadtxt(tmp_path / "SpikeTimes.txt", dtype=np.float64, ndmin=1)
    order = np.argsort(times, kind="stable")
    assert ids[order].tolist() == [0, 3, 1, 6, 7]
    assert np.allclose(times[order], [0.1, 0.2, 0.3, 0.5, 0.9], atol=1e-6)
    for f in ("PostsynapticIDs.txt", "SynapticWeights.txt"):
        assert (tmp_path / f).stat().st_size > 0
