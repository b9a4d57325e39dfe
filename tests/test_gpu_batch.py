"""Device-generated Brunel networks and batched persistent launches.

1. A network generated on the device equals, bit for bit, the FP32 oracle running the same
   network built on the host from the oracle's restatement of the generation rule.
2. Networks run in one batch give exactly what they give when run one at a time (the
   batch only changes which workgroups run which partitions), also when a few workgroups
   must run many networks one after another.
3. Batched FP32 rates are within 2 % of the FP64 oracle's.
4. The full sweep at its configured ranges (256 networks, N 1250-12500, 1 s): batched
   equals one-network-at-a-time bit for bit, and the FP32 rates averaged over 16
   realisations of the external drive lie within 2 % of the FP64 oracle's averages
   (tests/golden/sweep_fp64_rates.json).
"""
import json
import os

import numpy as np
import pytest

import oracle_py
from sweep_replay import replay_desc as host_replay, oracle_rate
from spike_amd import abi, sweep

pytestmark = pytest.mark.gpu
F32 = np.float32


def small_points(count, seed=7, n_range=(300, 1500)):
    return sweep.sample_points(count, seed=seed, n_range=n_range)


@pytest.mark.parametrize("partitions", [0, 5])
def test_generated_network_matches_oracle(partitions):
    p = sweep.SweepPoint(n=1000, g=5.0, nu_ratio=2.0, J=0.1e-3, delay_steps=12, seed=33)
    ng = 150
    print(f"[batch] generated N=1000 P={partitions}", flush=True)
    gpu = sweep.build(p, partitions)
    gpu.stats_begin(10, ng * p.delay_steps)
    gpu.run(0.0, ng, plasticity=False)
    d = host_replay(p)
    ora = oracle_py.OracleNetwork(d, precision=32)
    ora.run(0.0, ng)
    gi = gpu.info()
    assert (gi.frac_bits, gi.ring_slots, gi.history_words) == (ora.info()["frac"], ora.info()["B"], ora.info()["W"])
    assert np.array_equal(gpu.membrane().astype(np.float64), ora.membrane())
    assert np.array_equal(gpu.last_spike(0), ora.last_spike(0))
    assert np.array_equal(gpu.history(0), ora.history(0))
    assert np.array_equal(gpu.ring(), ora.ring())
    c, oc = gpu.counters(), ora.counters()
    assert c.spikes_neurons == oc["spikes_neurons"] > 0
    assert c.synaptic_events == oc["synaptic_events"] and c.external_events == oc["external_events"]
    s = gpu.stats()
    oi, ot = ora.take_spikes(0)
    assert s.spikes == len(oi)


@pytest.mark.parametrize("realisation", [1, 5])
def test_drive_realisation_matches_oracle(realisation):
    """Another realisation of the external drive (spk_set_seed after generation) keeps the
    connectivity and equals the oracle replaying the point with that drive seed."""
    p = sweep.SweepPoint(n=900, g=4.0, nu_ratio=2.5, J=0.1e-3, delay_steps=15, seed=77)
    ng = 120
    print(f"[batch] drive realisation {realisation}", flush=True)
    gpu = sweep.build(p, 0, realisation=realisation)
    gpu.run(0.0, ng, plasticity=False)
    d = host_replay(p)
    d.seed = sweep.drive_seed(p, realisation)
    ora = oracle_py.OracleNetwork(d, precision=32)
    ora.run(0.0, ng)
    assert np.array_equal(gpu.membrane().astype(np.float64), ora.membrane())
    assert np.array_equal(gpu.ring(), ora.ring())
    c, oc = gpu.counters(), ora.counters()
    assert c.spikes_neurons == oc["spikes_neurons"] > 0 and c.external_events == oc["external_events"]
    gpu.close()
    base = sweep.build(p, 0)
    base.run(0.0, ng, plasticity=False)
    assert base.counters().external_events != c.external_events
    base.close()


def _individual(points, ng_of, partitions_of):
    out = []
    for p in points:
        n = sweep.build(p, partitions_of(p))
        ng = ng_of(p)
        n.stats_begin(10, ng * p.delay_steps)
        n.run(0.0, ng, plasticity=False)
        s = n.stats()
        out.append((s.spikes, s.isi_cv, s.synchrony, n.membrane().copy()))
        n.close()
    return out


@pytest.mark.parametrize("workgroups", [0, 6])
def test_batch_equals_individual_runs(workgroups):
    pts = small_points(7, n_range=(300, 600))     # at most 5 workgroups per network
    ng_of = lambda p: 80 + p.delay_steps          # uneven run lengths
    part_of = lambda p: max(1, -(-p.n // 128))
    print(f"[batch] 7 networks, workgroups={workgroups}", flush=True)
    ref = _individual(pts, ng_of, part_of)
    nets = [sweep.build(p, part_of(p)) for p in pts]
    for n, p in zip(nets, pts):
        n.stats_begin(10, ng_of(p) * p.delay_steps)
    b = abi.Batch(nets, workgroups)
    if workgroups:
        assert b.workgroups == workgroups
    ms, est = b.run(np.zeros(len(pts)), [ng_of(p) for p in pts])
    assert ms > 0
    for k, (n, p) in enumerate(zip(nets, pts)):
        s = n.stats()
        assert (s.spikes, s.isi_cv, s.synchrony) == ref[k][:3], f"network {k}"
        assert np.array_equal(n.membrane(), ref[k][3]), f"network {k} membrane"
    # a second batch run continues every network where it stopped
    starts = [float(F32(F32(ng_of(p) * p.delay_steps) * F32(p.dt))) for p in pts]
    b.run(np.array(starts), [10] * len(pts))
    b.close()
    for n in nets:
        n.close()


def test_batch_graph_replays_equal_direct_launches():
    """A batch run is a captured hipGraph replayed by later runs.  Two copies of the same
    networks, one batch launching directly and one through its graph, run three times in a
    row (continuing, uneven group counts) with the host's own work between the runs; every
    run of the graph batch equals the direct one bit for bit."""
    pts = small_points(5, seed=3, n_range=(300, 700))
    part_of = lambda p: max(1, -(-p.n // 160))
    print("[batch] graph replays vs direct launches, 5 networks", flush=True)
    nets_d = [sweep.build(p, part_of(p)) for p in pts]
    nets_g = [sweep.build(p, part_of(p)) for p in pts]
    bd, bg = abi.Batch(nets_d, 0), abi.Batch(nets_g, 0)
    bd.set_graph(False)
    starts = np.zeros(len(pts))
    for k, ng in enumerate(([40] * 5, [25, 30, 35, 40, 45], [60] * 5)):
        for nd, ngr, p, g in zip(nets_d, nets_g, pts, ng):
            nd.stats_begin(10, g * p.delay_steps)
            ngr.stats_begin(10, g * p.delay_steps)
        bd.run(starts, ng)
        bg.run(starts, ng)
        for j, (nd, ngr) in enumerate(zip(nets_d, nets_g)):
            sd, sg = nd.stats(), ngr.stats()
            assert (sd.spikes, sd.isi_cv, sd.synchrony) == (sg.spikes, sg.isi_cv, sg.synchrony), f"run {k} net {j}"
            assert np.array_equal(nd.membrane(), ngr.membrane()), f"run {k} net {j} membrane"
            assert np.array_equal(nd.ring(), ngr.ring()), f"run {k} net {j} ring"
        starts = np.array([float(F32(F32(s) + F32(F32(g * p.delay_steps) * F32(p.dt))))
                           for s, g, p in zip(starts, ng, pts)])
    assert bg.graph_info() == (1, 2) and bd.graph_info() == (0, 0)
    for x in [bd, bg] + nets_d + nets_g:
        x.close()


def test_realisations_reuse_equals_fresh_builds():
    """sweep.run_realisations (one batch, drive seed set and state reset between runs, the
    graph replayed) gives exactly the rates of networks built afresh per realisation."""
    pts = small_points(4, seed=21, n_range=(400, 900))
    print("[batch] realisations through one batch vs fresh builds", flush=True)
    rates, ms, info = sweep.run_realisations(pts, 0.1, [0, 2, 5])
    assert info == (1, 2)
    for j, r in enumerate([0, 2, 5]):
        fresh = sweep.run_batched(pts, 0.1, realisation=r)
        assert np.array_equal(rates[:, j], fresh.rate), f"realisation {r}"
    assert not np.array_equal(rates[:, 0], rates[:, 1])


def test_batched_rates_within_2pc_of_fp64():
    pts = small_points(4, seed=11, n_range=(800, 1600))
    seconds = 1.0
    print(f"[batch] 4 networks 1 s FP32 batch vs FP64 oracle", flush=True)
    r = sweep.run_batched(pts, seconds, sweep.default_plan(pts, 128))
    for k, p in enumerate(pts):
        rate64, _ = oracle_rate(p, seconds, precision=64, threads=8)
        print(f"point {k}: N={p.n} rate FP32 batch {r.rate[k]:.3f} FP64 {rate64:.3f}", flush=True)
        assert abs(r.rate[k] - rate64) <= 0.02 * rate64


def test_every_partition_count_builds_co_resident_or_reports_lds():
    """Partition counts around the LDS limit: generation either fails with the LDS message
    (so the sweep builder splits further) or gives a network whose persistent grid is
    co-resident -- it runs alone and in a batch.  A partition sized just under the
    per-workgroup limit, with the kernel's static LDS on top, must not get through.  Delay 40
    steps (a 40-slot ring, G = 40: the LIF's chunked path with the inputs kept in the ring)
    puts the limit inside the range of partition counts probed."""
    p = sweep.SweepPoint(n=12000, g=5.0, nu_ratio=2.0, J=0.1e-3, delay_steps=40, seed=41)
    built, refused = 0, 0
    for P in range(12, 40):   # 12: at most 1000 neurons per partition
        print(f"[batch] N=12000 delay 40 P={P}", flush=True)
        try:
            net = abi.generated_brunel(sweep.brunel_params(p), P)
        except abi.SpikeError as e:
            assert "LDS" in str(e), str(e)
            refused += 1
            continue
        built += 1
        assert net.info().fused == 1
        net.run(0.0, 2, plasticity=False)
        b = abi.Batch([net], 0)
        b.run(np.zeros(1), [2])
        b.close()
        net.close()
    assert built > 0 and refused > 0


def test_full_sweep_at_configured_size():
    """The benchmark's sweep: 256 networks of N log-uniform in [1250, 12500], g U(3, 8),
    nu_ext/nu_thr U(0.9, 4), J U(0.05, 0.2) mV, delays 10-20 steps, 1 s each."""
    here = os.path.dirname(os.path.abspath(__file__))
    fx = json.load(open(os.path.join(here, "golden", "sweep_fp64_rates.json")))
    pts = sweep.sample_points(256)
    assert min(p.n for p in pts) >= 1250 and max(p.n for p in pts) <= 12500
    print("[batch] 256-network sweep: batched", flush=True)
    rb = sweep.run_batched(pts, 1.0, sweep.default_plan(pts, 384))
    print(f"[batch] batched {rb.ms:.1f} ms; plain", flush=True)
    rp = sweep.run_plain(pts, 1.0)
    print(f"[batch] plain {rp.ms:.1f} ms", flush=True)
    assert np.array_equal(rb.rate, rp.rate) and np.array_equal(rb.cv, rp.cv)
    assert np.array_equal(rb.sync, rp.sync)
    assert rb.ms < rp.ms
    idx = [e["index"] for e in fx["points"]]
    R = fx["realisations"]
    sub = [pts[k] for k in idx]
    r32 = np.zeros((len(idx), R))
    r32[:, 0] = rb.rate[idx]
    r32[:, 1:], _, info = sweep.run_realisations(sub, 1.0, range(1, R))
    assert info == (1, R - 2)
    m64 = np.array([np.mean(e["rate_fp64"]) for e in fx["points"]])
    err = np.abs(r32.mean(1) - m64) / m64
    for j, k in enumerate(idx):
        print(f"point {k} N={pts[k].n}: FP32 {r32[j].mean():.3f} Hz FP64 {m64[j]:.3f} Hz err {100 * err[j]:.2f} %")
    assert err.max() <= 0.02
