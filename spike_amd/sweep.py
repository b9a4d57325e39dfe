"""Parameter sweeps of Brunel networks: plain (one network per launch) and batched (all
networks in one persistent launch), with an autotuned workgroup-to-network tiling.

A sweep point is one Brunel (2000) network: N neurons (80 % excitatory), relative
inhibition g, external drive nu_ext/nu_thr, synaptic weight J and delay D steps; in-degree
rule C_E = eps N_E, C_I = eps N_I with eps = 0.1.  The network is generated on the device
(spk_create_brunel) and its statistics (rate, ISI CV, synchrony) are reduced in FP64 on the
device (spk_stats_*).
"""
from __future__ import annotations

import dataclasses
import math
import time
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import abi
from .netdesc import lif_constants

F32 = np.float32


@dataclasses.dataclass
class SweepPoint:
    n: int                 # neurons
    g: float               # relative inhibition
    nu_ratio: float        # nu_ext / nu_thr
    J: float               # EPSP amplitude (V)
    delay_steps: int       # axonal delay D (= timestep grouping G)
    seed: int
    eps: float = 0.1
    theta: float = 20e-3
    vreset: float = 10e-3
    tau: float = 20e-3
    tref: float = 2e-3
    dt: float = 1e-4

    @property
    def n_exc(self) -> int:
        return int(round(0.8 * self.n))

    @property
    def n_inh(self) -> int:
        return self.n - self.n_exc

    @property
    def ce(self) -> int:
        return max(1, int(self.eps * self.n_exc))

    @property
    def ci(self) -> int:
        return max(1, int(self.eps * self.n_inh))

    def groups(self, seconds: float) -> int:
        """SpikingModel::run's group count: ceil(T/dt) steps, truncated to whole groups."""
        steps = int(math.ceil(F32(seconds) / F32(self.dt)))
        return steps // self.delay_steps


def sample_points(count: int = 256, seed: int = 2024, n_range=(1250, 12500), g_range=(3.0, 8.0),
                  nu_range=(0.9, 4.0), J_range=(0.05e-3, 0.2e-3), delay_range=(10, 20)) -> List[SweepPoint]:
    """N log-uniform, g, nu_ext/nu_thr and J uniform, delay uniform in whole steps; per-point
    seeds = base seed + point index."""
    rng = np.random.default_rng(seed)
    n = np.exp(rng.uniform(np.log(n_range[0]), np.log(n_range[1]), count)).round().astype(int)
    g = rng.uniform(*g_range, count)
    nu = rng.uniform(*nu_range, count)
    J = rng.uniform(*J_range, count)
    d = rng.integers(delay_range[0], delay_range[1] + 1, count)
    return [SweepPoint(int(n[k]), float(g[k]), float(nu[k]), float(J[k]), int(d[k]), seed + k)
            for k in range(count)]


def point_constants(p: SweepPoint) -> Dict:
    """Every float32 constant of the network, computed as builders.brunel does (the same
    values go to the device generator and to a host replay of the network)."""
    cm = 200e-12
    g0 = cm / p.tau
    c, rp = lif_constants(1, cm, g0, p.dt)
    nu_thr = p.theta / (p.J * p.ce * p.tau)
    return dict(v0=F32(0.0), vreset=F32(p.vreset), threshold=F32(p.theta), decay_c=c[0], resist=rp[0],
                refractory=F32(p.tref), weight_exc=F32(p.J), weight_inh=F32(-p.g * p.J),
                ext_rate=F32(p.nu_ratio * nu_thr * p.ce), ext_weight=F32(p.J), dt=F32(p.dt))


def brunel_params(p: SweepPoint) -> abi.BrunelParams:
    k = point_constants(p)
    return abi.BrunelParams(
        n_exc=p.n_exc, n_inh=p.n_inh, ce=p.ce, ci=p.ci, delay_steps=p.delay_steps,
        weight_exc=k["weight_exc"], weight_inh=k["weight_inh"], v0=k["v0"], vreset=k["vreset"],
        threshold=k["threshold"], decay_c=k["decay_c"], resist=k["resist"],
        refractory_s=k["refractory"], ext_rate_hz=k["ext_rate"], ext_weight=k["ext_weight"],
        timestep=k["dt"], seed=p.seed)


# LDS of one workgroup when two share a compute unit (160 KB per CU on MI355X, less a
# margin for the kernel's static LDS)
LDS_HALF_CU = 80 * 1024 - 1024


DRIVE_SEED_STEP = 7919


def drive_seed(p: SweepPoint, realisation: int) -> int:
    """Seed of the external drive's realisation r of a point (r = 0: the point's own seed;
    the connectivity always comes from the point's seed)."""
    return int(p.seed) + DRIVE_SEED_STEP * int(realisation)


def build(p: SweepPoint, partitions: int = 0, lds_max: int = 0, realisation: int = 0) -> abi.Network:
    """The point's network on the device; with partitions > 0 the count is raised until the
    per-workgroup input ring fits in LDS (and in lds_max bytes when given).  `realisation`
    selects the external drive's seed (drive_seed)."""
    P = partitions
    while True:
        try:
            net = abi.generated_brunel(brunel_params(p), P)
        except abi.SpikeError as e:
            if "LDS" not in str(e) or P <= 0 or P >= p.n:
                raise
            P = min(p.n, int(P * 1.25) + 1)
            continue
        if lds_max <= 0 or P <= 0 or P >= p.n or net.info().lds_bytes <= lds_max:
            if realisation:
                abi.check(net.lib.spk_set_seed(net.h, drive_seed(p, realisation)))
            return net
        net.close()
        P = min(p.n, int(P * 1.25) + 1)


@dataclasses.dataclass
class SweepResult:
    ms: float                  # device time of the simulations (HIP events)
    wall_s: float              # host wall time including network generation
    rate: np.ndarray
    cv: np.ndarray
    sync: np.ndarray
    config: Dict


def _stats(nets, pts, seconds):
    rate, cv, sync = [], [], []
    for n in nets:
        s = n.stats()
        rate.append(s.rate_hz)
        cv.append(s.isi_cv)
        sync.append(s.synchrony)
    return np.array(rate), np.array(cv), np.array(sync)


def run_plain(points: Sequence[SweepPoint], seconds: float, bin_steps: int = 10) -> SweepResult:
    """One network at a time on the persistent kernel with its default partitioning."""
    t0 = time.perf_counter()
    ms = 0.0
    rate, cv, sync = [], [], []
    for p in points:
        net = build(p)
        ng = p.groups(seconds)
        net.stats_begin(bin_steps, ng * p.delay_steps)
        tm = net.run(0.0, ng, plasticity=False)
        ms += tm.total_ms
        s = net.stats()
        rate.append(s.rate_hz)
        cv.append(s.isi_cv)
        sync.append(s.synchrony)
        net.close()
    return SweepResult(ms, time.perf_counter() - t0, np.array(rate), np.array(cv), np.array(sync),
                       dict(mode="plain"))


def _build_all(points, partitions, seconds, bin_steps, profiling=False, lds_max=0, realisation=0):
    nets = []
    for p, P in zip(points, partitions):
        n = build(p, int(P), lds_max, realisation)
        n.stats_begin(bin_steps, p.groups(seconds) * p.delay_steps)
        if profiling:
            n.set_profiling(True)
        nets.append(n)
    return nets


def default_plan(points: Sequence[SweepPoint], neurons_per_wg: int = 384) -> Dict:
    """Tiling from sizes alone: ceil(N / neurons_per_wg) workgroups, largest first."""
    parts = [max(1, -(-p.n // neurons_per_wg)) for p in points]
    order = sorted(range(len(points)), key=lambda k: -points[k].n)
    return dict(partitions=parts, order=order, cost=None, name=f"size/{neurons_per_wg}",
                policy=dict(target=None, npw=int(neurons_per_wg)))


def run_batched(points: Sequence[SweepPoint], seconds: float, plan: Optional[Dict] = None,
                workgroups: int = 0, bin_steps: int = 10, profiling: bool = False,
                realisation: int = 0) -> SweepResult:
    """All networks in one persistent launch.  `plan` gives every network's workgroup count
    (partitions), the scheduling order and optionally the per-group cost the scheduler
    uses (microseconds); None uses default_plan.  `realisation` selects the external
    drive's realisation of every point (drive_seed)."""
    plan = plan or default_plan(points)
    t0 = time.perf_counter()
    order = list(plan["order"])
    nets = _build_all(points, plan["partitions"], seconds, bin_steps, profiling, plan.get("lds_max", 0),
                      realisation)
    batch = abi.Batch([nets[k] for k in order], workgroups)
    if plan.get("cost") is not None:
        batch.set_costs([plan["cost"][k] for k in order])
    ng = [points[k].groups(seconds) for k in order]
    ms, est = batch.run(np.zeros(len(order)), ng)
    rate, cv, sync = _stats(nets, points, seconds)
    extra = {}
    if profiling:
        ev = np.array([n.counters().synaptic_events for n in nets], dtype=np.float64)
        dur = np.array([(lambda pr: pr["s11"] - pr["s10"])(n.phase_profile()) for n in nets])
        extra = dict(events=ev, duration_us=dur)
    parts = [n.info().partitions for n in nets]
    wg, threads = batch.workgroups, batch.threads
    batch.close()
    for n in nets:
        n.close()
    r = SweepResult(ms, time.perf_counter() - t0, rate, cv, sync,
                    dict(mode="batched", plan=plan.get("name", "custom"), workgroups=wg, threads=threads,
                         estimated_ms=est, partitions_total=int(sum(parts))))
    r.extra = dict(extra, partitions=parts)
    return r


def run_realisations(points: Sequence[SweepPoint], seconds: float, realisations: Sequence[int],
                     plan: Optional[Dict] = None, workgroups: int = 0, bin_steps: int = 10):
    """The same networks under several realisations of the external drive (drive_seed):
    generated once, run as one batch, and for each realisation the drive seed set, the
    state reset and the batch run again -- every run after the first replays the batch's
    captured hipGraph.  Returns (rates [points][realisations], device ms per run,
    (graph captures, replays))."""
    plan = plan or default_plan(points)
    order = list(plan["order"])
    nets = _build_all(points, plan["partitions"], seconds, bin_steps, lds_max=plan.get("lds_max", 0))
    batch = abi.Batch([nets[k] for k in order], workgroups)
    if plan.get("cost") is not None:
        batch.set_costs([plan["cost"][k] for k in order])
    ng = [points[k].groups(seconds) for k in order]
    rates = np.zeros((len(points), len(realisations)))
    ms = []
    try:
        for j, r in enumerate(realisations):
            for p, n in zip(points, nets):
                abi.check(n.lib.spk_set_seed(n.h, drive_seed(p, r)))
                n.reset_state()
                n.stats_begin(bin_steps, p.groups(seconds) * p.delay_steps)
            m, _ = batch.run(np.zeros(len(order)), ng)
            ms.append(m)
            rates[:, j] = _stats(nets, points, seconds)[0]
        info = batch.graph_info()
    finally:
        batch.close()
        for n in nets:
            n.close()
    return rates, ms, info


def plan_from_work(points, events_per_group, target_events, neurons_per_wg_max, workgroups,
                   fixed_us, per_event_us, name=None) -> Dict:
    """Tiling from measured work: network k gets enough workgroups that each handles about
    `target_events` deliveries per group (at least ceil(N/neurons_per_wg_max), at most all
    workgroups); its per-group cost is fixed_us + per_event_us * events/workgroup; the
    longest networks are scheduled first."""
    K = len(points)
    parts, cost = [], []
    for k, p in enumerate(points):
        P = max(-(-p.n // neurons_per_wg_max), int(math.ceil(events_per_group[k] / max(1.0, target_events))))
        P = int(max(1, min(P, workgroups, p.n)))
        parts.append(P)
        cost.append(fixed_us + per_event_us * events_per_group[k] / P)
    total = [cost[k] * points[k].groups(1.0) * parts[k] for k in range(K)]
    order = sorted(range(K), key=lambda k: -cost[k] * points[k].groups(1.0))
    return dict(partitions=parts, order=order, cost=cost, name=name or f"work/{target_events:.0f}",
                lds_max=LDS_HALF_CU if neurons_per_wg_max <= 512 else 0,
                policy=dict(target=float(target_events), npw=int(neurons_per_wg_max)))


def autotune(points: Sequence[SweepPoint], probe_seconds: float = 0.02,
             targets=(2e4, 5e4, 1e5, 2e5, 5e5), neurons_per_wg_max=(384, 1024, 256), log=None,
             refine_seconds: float = 0.1, refine_top: int = 3, refine_repeats: int = 1,
             refine_samples: Optional[Sequence[Sequence[SweepPoint]]] = None) -> Dict:
    """The project's tiling autotuner.  (1) A probe run of the whole sweep with the
    size-only plan measures every network's deliveries per group and its time per group;
    (2) a least-squares fit of time = fixed + per_event * events/workgroup gives the cost
    model; (3) every candidate (events-per-workgroup target, most neurons per workgroup)
    is turned into a plan (plan_from_work) and timed on a probe run; (4) the refine_top
    fastest are timed again on refine_seconds, refine_repeats times each -- or, with
    refine_samples, once on each of those further samples of the distribution (each
    finalist's policy applied to the sample by plan_for) -- and the one with the lowest mean
    time is returned.  (A short probe ranks plans within a few per cent of each other
    unreliably: their schedules diverge later in the run; and one sample can favour a policy
    that another does not, by as much as the policies differ.)  Fewer, larger partitions pay the per-group fixed cost fewer times; a
    partition whose ring slice would not fit in LDS is split further when built.  Plans of
    at most 512 neurons per workgroup also keep every partition's LDS within half a compute
    unit's, so that the batch runs two 512-thread workgroups per compute unit."""
    base = default_plan(points, neurons_per_wg_max[0])
    r0 = run_batched(points, probe_seconds, base, profiling=True)
    groups = np.array([p.groups(probe_seconds) for p in points], dtype=np.float64)
    ev = r0.extra["events"] / np.maximum(groups, 1.0)
    us = r0.extra["duration_us"] / np.maximum(groups, 1.0)
    x = ev / np.array(r0.extra["partitions"], dtype=np.float64)
    A = np.stack([np.ones_like(x), x], axis=1)
    (fixed, per_ev), *_ = np.linalg.lstsq(A, us, rcond=None)
    fixed, per_ev = float(max(fixed, 1.0)), float(max(per_ev, 1e-6))
    wg = r0.config["workgroups"]
    trials = [dict(plan=base["name"], ms=r0.ms)]
    best = dict(plan=base, ms=r0.ms)
    timed = [(r0.ms, base)]
    if log:
        log(f"autotune probe {base['name']}: {r0.ms:.2f} ms; fit {fixed:.2f} us + {per_ev * 1e3:.3f} ns/event")
    for npw, tgt in [(npw, tgt) for npw in neurons_per_wg_max for tgt in targets]:
        plan = plan_from_work(points, ev, tgt, npw, wg, fixed, per_ev, name=f"work/{tgt:.0f}/{npw}")
        try:
            r = run_batched(points, probe_seconds, plan)
        except abi.SpikeError as e:
            trials.append(dict(plan=plan["name"], error=str(e)))
            continue
        trials.append(dict(plan=plan["name"], ms=r.ms, estimated_ms=r.config["estimated_ms"]))
        timed.append((r.ms, plan))
        if log:
            log(f"autotune {plan['name']}: {r.ms:.2f} ms (scheduler estimate {r.config['estimated_ms']:.2f}, "
                f"{r.config['workgroups']} x {r.config['threads']} threads)")
        if r.ms < best["ms"]:
            best = dict(plan=plan, ms=r.ms)
    if refine_top > 1 and refine_seconds > probe_seconds:
        finalists = [pl for _, pl in sorted(timed, key=lambda x: x[0])[:refine_top]]
        best = None
        for plan in finalists:
            if refine_samples:
                runs = []
                for smp in refine_samples:
                    sp, pms = plan_for(smp, plan.get("policy"))
                    runs.append(run_batched(smp, refine_seconds, sp).ms + pms)
            else:
                runs = [run_batched(points, refine_seconds, plan).ms for _ in range(max(1, refine_repeats))]
            ms = float(np.mean(runs))
            trials.append(dict(plan=plan["name"], refine_ms=ms, refine_runs=runs))
            if log:
                log(f"autotune refine {plan['name']}: {ms:.2f} ms over {refine_seconds} s "
                    f"(mean of {len(runs)}: {', '.join(f'{x:.1f}' for x in runs)})")
            if best is None or ms < best["ms"]:
                best = dict(plan=plan, ms=ms)
    best["trials"] = trials
    best["fit"] = dict(fixed_us=fixed, per_event_us=per_ev)
    best["policy"] = best["plan"].get("policy")
    return best


def plan_for(points: Sequence[SweepPoint], policy: Optional[Dict], probe_seconds: float = 0.02):
    """The tiling of a tuned policy (autotune(...)["policy"], tuned on other points) for
    `points`: a probe run with the size-only plan measures their deliveries per group and the
    cost fit, then plan_from_work applies the policy.  Returns (plan, probe device ms)."""
    npw = policy["npw"] if policy else 384
    base = default_plan(points, npw)
    if not policy or policy.get("target") is None:
        return base, 0.0            # a size-only tiling needs no probe
    r0 = run_batched(points, probe_seconds, base, profiling=True)
    groups = np.array([p.groups(probe_seconds) for p in points], dtype=np.float64)
    ev = r0.extra["events"] / np.maximum(groups, 1.0)
    us = r0.extra["duration_us"] / np.maximum(groups, 1.0)
    x = ev / np.array(r0.extra["partitions"], dtype=np.float64)
    A = np.stack([np.ones_like(x), x], axis=1)
    (fixed, per_ev), *_ = np.linalg.lstsq(A, us, rcond=None)
    fixed, per_ev = float(max(fixed, 1.0)), float(max(per_ev, 1e-6))
    plan = plan_from_work(points, ev, policy["target"], policy["npw"], r0.config["workgroups"], fixed, per_ev,
                          name=f"work/{policy['target']:.0f}/{policy['npw']}")
    return plan, r0.ms
