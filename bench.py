#!/usr/bin/env python3
"""Brunel10K benchmark: synaptic events per second on one MI355X.

Workload (the Brunel balanced network): 8000 excitatory + 2000 inhibitory LIF neurons
(tau_m 20 ms, theta 20 mV, V_reset 10 mV, 2 ms refractory), delta synapses J = 0.1 mV,
g = 5, fixed in-degree C_E = 800 / C_I = 200, delay 1.5 ms (G = 15 steps), external
Poisson drive nu_ext = 2 nu_thr per neuron (20 kHz x J), dt = 0.1 ms, FP32, synthetic
connectivity (seed 42).  One bench step = one simulated second (666 groups x 15 steps,
the reference's truncated group count).  Synaptic events = recurrent deliveries
(spikes x out-degree) + external Poisson events.

Second quantity of the metric ("sweep" in the JSON line): 256 Brunel networks of uneven
size, 1 s each, in one batched persistent launch versus one network per launch, with the
batched FP32 rates checked against the FP64 oracle on a sample of the points.  Also
reported: the surrogate MLP kernels ("mlp") and 2 s of the large plastic network ("large").

Prints one JSON line with the metric, the roofline of the dominant kernel and the CPU
baseline (the FP32 oracle on the host cores, on a bounded sample of the same network).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(n_neurons, steps, spikes, rec_events, b_t=4, b_w=4):
    """Algorithmic HBM bytes of the hot path, per unit of work:
    per neuron per step 24 B (V r/w 8, refractory state r/w 8, ring slot read+clear 8);
    per spike 16 B (list write 4, list read 4, row pointers 8);
    per recurrent delivery b_t = 4 B target + b_w = 4 B FP32 weight.
    External Poisson events are generated in registers (0 B).  These are the algorithm's
    bytes at FP32, whatever record format the kernel reads: a kernel that reads 2-byte
    compact records moves fewer bytes than this (see the measured traffic)."""
    return n_neurons * steps * 24 + spikes * 16 + rec_events * (b_t + b_w)


def measured_traffic(form, path):
    """HBM bytes per launch of the fused kernel, from the committed rocprofv3 PMC summary
    (FETCH_SIZE and WRITE_SIZE, each scaled by the factor a 1 GiB copy gives); None when
    the summary for this workload is absent."""
    if form != "config" or path != "fused":
        return None
    f = os.path.join(ROOT, "profiles", "k_fused_brunel10k_config.txt")
    if not os.path.exists(f):
        return None
    for line in open(f):
        if line.startswith("traffic_bytes_per_launch:"):
            return int(float(line.split(":")[1]))
    return None


def host_cpus():
    """Host CPU facts for the baseline: logical CPUs the OS reports
    (std::thread::hardware_concurrency's value), the CPUs this process may run on, and the
    physical cores lscpu counts."""
    logical = os.cpu_count() or 1
    try:
        allowed = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        allowed = logical
    physical = None
    try:
        import subprocess
        out = subprocess.run(["lscpu", "-p=Core,Socket"], capture_output=True, text=True, timeout=10).stdout
        physical = len({ln for ln in out.splitlines() if ln and not ln.startswith("#")})
    except (OSError, subprocess.SubprocessError):
        pass
    return logical, allowed, physical


def cpu_baseline(desc, warm_groups, sample_groups, one_groups, threads):
    """FP32 oracle on the host: the network is first run through the same warm-up the GPU
    gets (multithreaded), then a steady-state sample of `sample_groups` groups is timed with
    `threads` threads and one of `one_groups` groups with one thread.  Returns (events/s,
    seconds, events) for both."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py
    o = oracle_py.OracleNetwork(desc, precision=32, threads=threads)
    t = 0.0
    if warm_groups:
        o.run(t, warm_groups)
        t = float(desc.group_time(t, warm_groups))

    def timed(groups):
        nonlocal t
        c0 = o.counters()
        t0 = time.perf_counter()
        o.run(t, groups)
        dt = time.perf_counter() - t0
        t = float(desc.group_time(t, groups))
        c1 = o.counters()
        ev = (c1["synaptic_events"] + c1["external_events"]) - (c0["synaptic_events"] + c0["external_events"])
        return ev / dt, dt, ev

    multi = timed(sample_groups)
    o.set_threads(1)
    one = timed(one_groups)
    o.close()
    return multi, one


SWEEP_FIXTURE = os.path.join(ROOT, "tests", "golden", "sweep_fp64_rates.json")


def sweep_leg(n_points, seconds, cpu_points, threads, log, tune_seed=7):
    """Batched sweep: `n_points` Brunel networks (N log-uniform 1250-12500, g U(3,8),
    nu_ext/nu_thr U(0.9,4), J U(0.05,0.2) mV, delay U{10..20} steps, eps 0.1, seeds
    2024 + index) simulated for `seconds` each, (1) in one persistent batched launch and
    (2) one network per launch.  The batched tiling policy is tuned on a disjoint sample
    of the same distribution (seed `tune_seed`; its finalists timed over the full `seconds`
    on three such samples, seeds `tune_seed` .. `tune_seed` + 2, each with its own plan);
    the benchmark points then get their plan
    from a short probe run, whose device time counts in the batched time.  Points/s =
    points / device seconds (HIP events; network generation is reported separately).

    Rate gate (<= 2 % vs FP64): the chaotic 1 s rate of one run differs from another
    realisation of the external drive by up to ~2 % at low rates, so the gate compares
    means over R realisations of the drive (same connectivity): for every point of the FP64
    fixture (every 8th point, R = 16, tests/golden/sweep_fp64_rates.json written by the FP64
    oracle), err = |mean_R(batched FP32) - mean_R(FP64)| / mean_R(FP64), gate max err <= 2 %.
    The CPU baseline of this metric times the FP64 oracle on `cpu_points` of those points."""
    import json as _json
    import numpy as np
    from spike_amd import sweep
    pts = sweep.sample_points(n_points)
    t0 = time.perf_counter()
    best = sweep.autotune(sweep.sample_points(n_points, seed=tune_seed), log=log, refine_seconds=seconds,
                          refine_samples=[sweep.sample_points(n_points, seed=tune_seed + j) for j in range(3)])
    tune_s = time.perf_counter() - t0
    plan, probe_ms = sweep.plan_for(pts, best["policy"])
    rb = sweep.run_batched(pts, seconds, plan)
    batched_ms = rb.ms + probe_ms
    log(f"sweep batched: {rb.ms:.2f} ms + probe {probe_ms:.2f} ms ({rb.config})")
    rp = sweep.run_plain(pts, seconds)
    log(f"sweep plain: {rp.ms:.2f} ms")
    speedup = rp.ms / batched_ms
    out = {
        "points": n_points, "seconds_per_point": seconds,
        "batched_points_per_s": round(n_points / (batched_ms / 1e3), 2),
        "plain_points_per_s": round(n_points / (rp.ms / 1e3), 2),
        "speedup_vs_plain": round(speedup, 3),
        "gate_speedup": ">= 10", "gate_speedup_ok": bool(speedup >= 10.0),
        "batched_ms": round(batched_ms, 3), "batched_run_ms": round(rb.ms, 3), "probe_ms": round(probe_ms, 3),
        "plain_ms": round(rp.ms, 3),
        "plan": plan["name"], "tuned_on": f"sample_points({n_points}, seed={tune_seed}); finalists on seeds {tune_seed}-{tune_seed + 2}",
        "workgroups": rb.config["workgroups"], "threads_per_workgroup": rb.config["threads"],
        "autotune_s": round(tune_s, 2), "cost_fit_us": best["fit"],
        "batched_vs_plain_rate_max_rel_diff": float(np.max(np.abs(rb.rate - rp.rate) / np.maximum(rp.rate, 1e-12))),
        "mean_rate_hz": round(float(rb.rate.mean()), 3),
    }
    if os.path.exists(SWEEP_FIXTURE):
        fx = _json.load(open(SWEEP_FIXTURE))
        if fx["points_in_sweep"] == n_points and fx["seconds"] == seconds:
            idx = [e["index"] for e in fx["points"]]
            R = int(fx["realisations"])
            r64 = np.array([e["rate_fp64"] for e in fx["points"]], dtype=np.float64)
            sub = [pts[k] for k in idx]
            r32 = np.zeros((len(idx), R))
            r32[:, 0] = rb.rate[idx]
            # one batch for the other realisations: its runs after the first replay its hipGraph
            r32[:, 1:], real_ms, ginfo = sweep.run_realisations(sub, seconds, range(1, R))
            out["realisation_runs"] = {"runs": R - 1, "graph_captures": ginfo[0], "graph_replays": ginfo[1],
                                       "device_ms_mean": round(float(np.mean(real_ms)), 3)}
            m32, m64 = r32.mean(1), r64.mean(1)
            err = np.abs(m32 - m64) / m64
            se = np.sqrt(r32.var(1, ddof=1) / R + r64.var(1, ddof=1) / R) / m64
            for j, k in enumerate(idx):
                log(f"sweep point {k} N={pts[k].n}: batched FP32 mean {m32[j]:.3f} Hz, FP64 mean {m64[j]:.3f} Hz "
                    f"over {R} drive realisations, err {100 * err[j]:.2f} % (noise {100 * se[j]:.2f} %)")
            out.update({
                "rate_gate": f"max over {len(idx)} points of |mean of {R} FP32 realisations - mean of {R} FP64 "
                             f"realisations| / FP64 mean <= 2 %",
                "rate_err_vs_fp64_max": round(float(err.max()), 6),
                "rate_err_vs_fp64_mean": round(float(err.mean()), 6),
                "rate_err_noise_max": round(float(se.max()), 6),
                "gate_rate_ok": bool(err.max() <= 0.02),
                "rate_err_points": idx,
            })
            if cpu_points > 0:
                sys.path.insert(0, os.path.join(ROOT, "oracle"))
                from sweep_replay import oracle_rate
                secs = 0.0
                # the middle point of each size quantile of the fixture's points
                step = max(1, len(idx) // cpu_points)
                cp = sorted(idx, key=lambda k: pts[k].n)[step // 2::step][:cpu_points]
                for k in cp:
                    _, el = oracle_rate(pts[k], seconds, precision=64, threads=threads)
                    secs += el
                out["cpu_baseline"] = {"value": round(len(cp) / secs, 4), "unit": "sweep points/s",
                                       "cores": threads, "kind": "port",
                                       "sample": f"FP64 oracle, {threads} threads, on points {cp} of the "
                                                 f"{n_points}, {seconds} s each, {secs:.2f} s"}
    return out


MFMA_BF16_PEAK_TFLOPS = 2500.0


def mlp_leg(log, members=8, rows=1 << 20, batch=4096, steps=200):
    """Surrogate MLP kernels (5 -> 4 x 256 SiLU -> 3, BF16 MFMA, FP32 accumulation):
    inference of `rows` rows by `members` ensemble members in one launch, and training
    steps (forward, backward, weight gradient, Adam) on batches of `batch` rows, replayed as
    a HIP graph.  Algorithmic FLOPs per row per member: 2 (5*256 + 3*256*256 + 256*3);
    a training step counts three times that per row (forward, input gradients, weight
    gradients).  Times are HIP events on the surrogate's stream."""
    import numpy as np
    from spike_amd import mlp
    fl = 2.0 * (5 * 256 + 3 * 256 * 256 + 256 * 3)
    rng = np.random.default_rng(3)
    m = mlp.Mlp(members, batch)
    m.set_params(mlp.init_params(4, members))
    x = rng.uniform(-1.0, 1.0, (rows, 5)).astype(np.float32)
    ms = []
    for _ in range(5):
        m.predict(x)
        ms.append(m.timing().forward_ms)
    inf_ms = float(np.mean(ms[1:]))
    inf_tflops = members * rows * fl / (inf_ms / 1e3) / 1e12
    data_x = rng.uniform(-1.0, 1.0, (50000, 5)).astype(np.float32)
    data_t = rng.standard_normal((50000, 3)).astype(np.float32)
    m.set_data(data_x, data_t)
    m.gradients(batch)
    k = m.timing()
    m.train(20, batch, 1e-3, 1e-5, 10 * steps, use_graph=True, record_loss=False)
    m.train(steps, batch, 1e-3, 1e-5, 10 * steps, use_graph=True, record_loss=False)
    step_ms = m.timing().total_ms / steps
    train_tflops = 3 * members * batch * fl / (step_ms / 1e3) / 1e12
    m.close()
    log(f"mlp: inference {inf_ms:.3f} ms ({inf_tflops:.0f} TFLOP/s), training step {step_ms * 1e3:.1f} us "
        f"({train_tflops:.0f} TFLOP/s)")
    return {
        "workload": f"{members} members x {rows} rows inference; training batch {batch} x {members} members",
        "dtype": "bf16 (fp32 accumulate)",
        "inference_ms": round(inf_ms, 4), "inference_rows_per_s": round(members * rows / (inf_ms / 1e3), 1),
        "roofline": {"bound": "mfma", "kernel": "k_mlp_forward", "achieved": round(inf_tflops, 1),
                     "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(inf_tflops / MFMA_BF16_PEAK_TFLOPS, 4), "traffic": None},
        "train_step_us": round(step_ms * 1e3, 2), "train_tflops": round(train_tflops, 1),
        "train_kernels_us": {"forward": round(k.forward_ms * 1e3, 2), "backward": round(k.backward_ms * 1e3, 2),
                             "wgrad": round(k.wgrad_ms * 1e3, 2), "adam": round(k.update_ms * 1e3, 2)},
    }


# Algorithmic bytes of the large plastic network's step (per-phase kernels), per item:
#   neuron   56 B: V read+write 8, v0/vreset/threshold/decay/resistance 20, last spike 4,
#                  its ring slot read+clear 16, history word read+write 8
#   input    12 B: last spike 4, history word read+write 8
#   delivery 24 B: target 4, FP16 weight 2, delay 2, scale 4, 64-bit ring add (read+write) 12
#   plastic update 24 B: per synapse of a spiking excitatory neuron's rule lists (800 out, 800 in
#                  on average): its index or position 4, the other side's last spike 4,
#                  weight + Kahan word read+write 8, the other side's trace 4, source unit 4
LARGE_BYTES = dict(neuron=56, input=12, delivery=24, plastic=24)


def large_leg(seconds, log):
    """Large plastic network at full size: 5e5 LIF neurons (80 % excitatory), 1000 random
    outgoing synapses each (5e8 synapses, delays 1..20 steps so G = 1, FP16 weights),
    2.5e4 Poisson inputs at 10 Hz with 1000 targets each, nearest-spike weight-dependent
    STDP on the 3.2e8 E->E synapses; dt = 0.1 ms.  Synapses generated on the device by the
    Philox fixed out-degree rule (seed 42); per-phase kernels (the ring slice of a
    partition does not fit in LDS), weight-dependent rule driven by spikes.  Times are HIP
    events on the network's stream, after 20 warm-up steps (the first runs the rule's
    per-synapse form once)."""
    from spike_amd import abi, builders
    from spike_amd.netdesc import WEIGHTS_FP16
    t0 = time.perf_counter()
    d = builders.large_plastic(monitors=())
    d.weight_format = WEIGHTS_FP16
    net = abi.Network(d)
    build_s = time.perf_counter() - t0
    n, n_in, S = d.n_neurons, d.n_inputs, d.n_synapses
    P = net.rule_size(0)
    o = d.outdegree
    warm = 20
    t = float(d.group_time(0.0, warm))
    groups = d.groups_for(seconds)
    log(f"large: synapses generated and prepared on the device in {build_s:.2f} s")
    net.run(0.0, warm)
    tm = net.run(t, groups)
    net.close()
    steps = groups   # G = 1
    ms = tm.total_ms
    # plastic updates: each excitatory spike visits its rule synapses as source and as target
    plastic_per_spike = 2.0 * P / o.n_exc
    exc_spikes = tm.spikes_neurons * o.n_exc / n
    B = LARGE_BYTES
    algo = (steps * (n * B["neuron"] + n_in * B["input"]) + tm.synaptic_events * B["delivery"]
            + exc_spikes * plastic_per_spike * B["plastic"])
    achieved = algo / (ms / 1e3) / 1e9
    log(f"large: {seconds} s simulated in {ms:.1f} ms ({ms / steps * 1e3:.1f} us per step), "
        f"{tm.spikes_neurons} spikes, {tm.synaptic_events} deliveries, {achieved:.0f} GB/s algorithmic")
    return {
        "workload": "large plastic net: 5e5 LIF (80/20), 1000 random out-synapses each, delays 1-20 steps, "
                    "FP16 weights, 2.5e4 Poisson inputs at 10 Hz, nearest-spike WD STDP on E->E, dt 0.1 ms",
        "n_neurons": n, "n_inputs": n_in, "n_synapses": S, "plastic_synapses": P,
        "simulated_s": seconds, "steps": steps, "ms": round(ms, 2),
        "us_per_step": round(ms / steps * 1e3, 2),
        "syn_events_per_s": round(tm.synaptic_events / (ms / 1e3), 1),
        "rate_hz": round(tm.spikes_neurons / n / (steps * 1e-4), 3),
        "spikes": tm.spikes_neurons, "input_spikes": tm.spikes_inputs, "deliveries": tm.synaptic_events,
        "build_s": round(build_s, 2), "builder": "device (Philox fixed out-degree)", "path": "per-phase kernels",
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "bytes_model": B, "scope": "whole step (both launches), algorithmic bytes"},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--form", default="config", choices=["config", "reference"])
    ap.add_argument("--nu-ext-ratio", type=float, default=2.0,
                    help="external drive nu_ext / nu_thr of the Brunel network (0: no drive, no spikes: the "
                         "persistent kernel's per-group cost with nothing to deliver)")
    ap.add_argument("--path", default="fused", choices=["fused", "launches"])
    ap.add_argument("--cpu-groups", type=int, default=2000,
                    help="timestep groups of the CPU-baseline sample (15 steps each), after the warm-up")
    ap.add_argument("--cpu-one-groups", type=int, default=200,
                    help="timestep groups of the one-thread CPU-baseline sample")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--sweep-points", type=int, default=256,
                    help="Brunel networks in the batched-sweep measurement (0 skips it)")
    ap.add_argument("--sweep-seconds", type=float, default=1.0)
    ap.add_argument("--sweep-cpu-sample", type=int, default=4,
                    help="sweep points timed on the FP64 oracle (the sweep's CPU baseline)")
    ap.add_argument("--no-mlp", action="store_true", help="skip the surrogate MLP kernels")
    ap.add_argument("--large-seconds", type=float, default=2.0,
                    help="simulated seconds of the large plastic network (0: skip it)")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="also run a plain device copy of this many MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(the per-window spike exchange for neuron-range sharding is designed, not built)",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from spike_amd import abi, builders

    if args.calib_copy_mib:
        # A plain copy kernel of known size: rocprof's FETCH_SIZE / WRITE_SIZE for it give
        # the factor between the counters and the bytes a kernel really moves.
        import torch
        n = args.calib_copy_mib * (1 << 20) // 4
        x = torch.ones(n, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        del x, y

    desc = builders.brunel(seed=42, form=args.form, nu_ext_ratio=args.nu_ext_ratio)
    G = desc.timestep_grouping()
    groups = desc.groups_for(1.0)
    net = abi.Network(desc)
    info = net.info()
    force = args.path == "launches"
    # warmup: run the network into its steady state
    t = 0.0
    for _ in range(args.warmup):
        net.run(t, groups, plasticity=False, force_launches=force)
        t = float(desc.group_time(t, groups))
    tot_ms = 0.0
    kern_ms = 0.0
    launches = 0
    events = 0
    rec = 0
    spikes = 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        tm = net.run(t, groups, plasticity=False, force_launches=force)
        t = float(desc.group_time(t, groups))
        tot_ms += tm.total_ms
        kern_ms += tm.kernel_ms
        launches += tm.kernel_launches
        events += tm.synaptic_events + tm.external_events
        rec += tm.synaptic_events
        spikes += tm.spikes_neurons + tm.spikes_inputs
    wall = time.perf_counter() - t0
    net.close()
    steps_sim = groups * G * args.steps
    value = events / (tot_ms / 1e3)
    rate_hz = spikes / desc.n_neurons / (steps_sim * desc.dt)
    # roofline of the dominant kernel (the fused group kernel)
    bytes_alg = algorithmic_bytes(desc.n_neurons, steps_sim, spikes, rec)
    if launches > 0 and kern_ms > 0:
        avg_launch_ms = kern_ms / launches
        per_launch_bytes = bytes_alg / launches
        achieved = per_launch_bytes / (avg_launch_ms / 1e3) / 1e9
    else:
        avg_launch_ms = tot_ms / max(1, args.steps)
        achieved = bytes_alg / (tot_ms / 1e3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                "traffic": measured_traffic(args.form, args.path),
                "avg_launch_ms": round(avg_launch_ms, 5)}
    cpu = None
    if not args.no_cpu:
        logical, allowed, physical = host_cpus()
        threads = args.cpu_threads or allowed
        (v, secs, ev), (v1, secs1, ev1) = cpu_baseline(desc, groups * args.warmup, args.cpu_groups,
                                                        args.cpu_one_groups, threads)
        cpu = {"value": round(v, 1), "unit": "syn events/s", "cores": threads, "kind": "port",
               "threads": threads, "hardware_concurrency": logical, "allowed_cpus": allowed,
               "physical_cores": physical, "one_thread_value": round(v1, 1),
               "sample": f"FP32 oracle, Brunel10K {args.form} form, after the GPU's warm-up ({args.warmup} "
                         f"simulated s = {groups * args.warmup} groups): {args.cpu_groups} groups "
                         f"({args.cpu_groups * G} steps) with {threads} threads in {secs:.2f} s, {ev} events; "
                         f"{args.cpu_one_groups} groups with 1 thread in {secs1:.2f} s, {ev1} events"}
    out = {
        "metric": "syn events/s (Brunel10K); batched sweep points/s vs plain at \u22642% rate err",
        "value": round(value, 1),
        "unit": "syn events/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(tot_ms / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (Brunel network, seed 42)",
        "config": {"workload": f"Brunel10K {args.form} form: 8000E+2000I LIF, C_E=800/C_I=200, "
                               f"J=0.1 mV g=5, delay 1.5 ms, nu_ext={args.nu_ext_ratio:g}nu_thr, dt 0.1 ms, 1 s per step",
                   "n_neurons": desc.n_neurons, "n_synapses": desc.n_synapses,
                   "timestep_grouping": G, "groups_per_step": groups, "path": args.path,
                   "record_bytes": info.record_bytes if args.path == "fused" else 8},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "rate_hz": round(rate_hz, 3),
        "events_per_step": events // max(1, args.steps),
        "recurrent_events_per_step": rec // max(1, args.steps),
        "spikes_per_step": spikes // max(1, args.steps),
        "algorithmic_bytes_per_step": int(bytes_alg) // max(1, args.steps),
        "wall_s": round(wall, 3),
    }
    if cpu:
        out["speedup_vs_cpu"] = round(value / cpu["value"], 2)
    if args.sweep_points > 0:
        log = lambda m: print(m, file=sys.stderr, flush=True)
        threads = args.cpu_threads or host_cpus()[1]
        out["sweep"] = sweep_leg(args.sweep_points, args.sweep_seconds,
                                 0 if args.no_cpu else args.sweep_cpu_sample, threads, log)
    if not args.no_mlp:
        out["mlp"] = mlp_leg(lambda m: print(m, file=sys.stderr, flush=True))
    if args.large_seconds > 0:
        out["large"] = large_leg(args.large_seconds, lambda m: print(m, file=sys.stderr, flush=True))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
