"""Host replay of a device-generated Brunel sweep point, for the CPU oracle.

TEST INFRASTRUCTURE ONLY (the checker of spk_create_brunel and of batched sweeps): used by
tests/ and by bench.py's CPU-baseline leg.  The connectivity comes from the oracle's own
restatement of the device generator (or_brunel_connectivity), every float32 constant from
spike_amd.sweep.point_constants, so the replayed network is the point's network.
"""
from __future__ import annotations

import time

import numpy as np

import oracle_py
from spike_amd import sweep
from spike_amd.netdesc import NetDesc, SYN_VOLTAGE

F32 = np.float32


def replay_desc(p: sweep.SweepPoint, monitors=(0,)) -> NetDesc:
    k = sweep.point_constants(p)
    N = p.n
    pre = oracle_py.brunel_connectivity(p.seed, p.n_exc, p.n_inh, p.ce, p.ci)
    C = p.ce + p.ci
    post = np.repeat(np.arange(N, dtype=np.int32), C)
    w = np.where(np.arange(C) < p.ce, k["weight_exc"], k["weight_inh"]).astype(F32)
    return NetDesc(n_neurons=N, n_inputs=0, kind=SYN_VOLTAGE, dt=float(k["dt"]), seed=p.seed,
                   v0=np.full(N, k["v0"], F32), vreset=np.full(N, k["vreset"], F32),
                   threshold=np.full(N, k["threshold"], F32), decay_c=np.full(N, k["decay_c"], F32),
                   resist=np.full(N, k["resist"], F32), refractory=float(k["refractory"]), background=0.0,
                   ext_rate=float(k["ext_rate"]), ext_weight=float(k["ext_weight"]),
                   pre=pre.reshape(-1), post=post, weight=np.tile(w, N),
                   delay=np.full(N * C, p.delay_steps, np.int32), monitors=list(monitors), name="replay")


def oracle_rate(p: sweep.SweepPoint, seconds: float, precision: int = 64, threads: int = 8,
                drive_seed=None):
    """Population rate (Hz) of the point over `seconds` on the oracle, and the seconds the
    oracle's run loop took (network construction excluded).  `drive_seed` replaces the
    seed of the external Poisson drive only (same connectivity): two FP64 runs that differ
    in it measure how much the rate of a 1 s window varies from one realisation of the
    drive to another."""
    d = replay_desc(p, monitors=())
    if drive_seed is not None:
        d.seed = int(drive_seed)
    o = oracle_py.OracleNetwork(d, precision=precision, threads=threads)
    ng = p.groups(seconds)
    t0 = time.perf_counter()
    o.run(0.0, ng)
    el = time.perf_counter() - t0
    c = o.counters()
    o.close()
    rate = c["spikes_neurons"] / (p.n * ng * p.delay_steps * float(F32(p.dt)))
    return rate, el
