"""Synthetic networks of the benchmark configurations, as NetDesc objects.

Connectivity here is drawn with numpy's PCG64 from a fixed seed, or, for the large plastic
network, by the library's Philox fixed out-degree generator on the device (the C++
front-end keeps the reference's glibc rand() construction order; these builders are for
tests and the benchmark, where only the network's statistics matter).
"""
from __future__ import annotations

import numpy as np

from .netdesc import (F32, NetDesc, OutDegree, Rule, RULE_EVANS, RULE_VOGELS, RULE_WD, SYN_CONDUCTANCE, SYN_CURRENT,
                      SYN_VOLTAGE, conductance_decay, lif_constants)


def _lif(d: NetDesc, n, v0, vreset, theta, cm=200e-12, g0=10e-9, tref=2e-3, ibg=0.0):
    d.v0 = np.full(n, F32(v0), F32)
    d.vreset = np.full(n, F32(vreset), F32)
    d.threshold = np.full(n, F32(theta), F32)
    d.decay_c, d.resist = lif_constants(n, cm, g0, d.dt)
    d.refractory = float(F32(tref))
    d.background = float(F32(ibg))


def brunel(n_exc=8000, n_inh=2000, eps=0.1, J=0.1e-3, g=5.0, delay_steps=15, theta=20e-3,
           vreset=10e-3, tau=20e-3, tref=2e-3, nu_ext_ratio=2.0, seed=42, dt=1e-4,
           form="config", monitors=(), reference_input_rate=20.0) -> NetDesc:
    """Brunel (2000) balanced network of delta-current LIF neurons.

    form="config": per-neuron external Poisson drive nu_ext = ratio * nu_thr * C_E
    generated in registers (1e7 recurrent synapses at the default size).
    form="reference": the reference example's encoding (Examples/Brunel10K.cpp:137-195):
    10^4 shared Poisson inputs at 20 Hz, each neuron receiving eps*10^4 of them.
    """
    rng = np.random.default_rng(seed)
    N = n_exc + n_inh
    ce, ci = int(eps * n_exc), int(eps * n_inh)
    n_in = 10000 if form == "reference" else 0
    d = NetDesc(n_neurons=N, n_inputs=n_in, kind=SYN_VOLTAGE, dt=dt, seed=seed, name=f"brunel_{form}")
    cm, g0 = 200e-12, 200e-12 / tau
    _lif(d, N, 0.0, vreset, theta, cm=cm, g0=g0, tref=tref)
    pre_e = rng.integers(0, n_exc, size=(N, ce), dtype=np.int64)
    pre_i = rng.integers(n_exc, N, size=(N, ci), dtype=np.int64)
    post = np.repeat(np.arange(N, dtype=np.int32), ce + ci)
    pre = np.concatenate([pre_e, pre_i], axis=1).reshape(-1).astype(np.int32)
    w = np.where(pre < n_exc, F32(J), F32(-g * J)).astype(F32)
    parts_pre, parts_post, parts_w = [pre], [post], [w]
    if form == "reference":
        k = int(eps * n_in)
        ipre = rng.integers(0, n_in, size=(N, k), dtype=np.int64)
        parts_pre.append((-ipre.reshape(-1) - 1).astype(np.int32))
        parts_post.append(np.repeat(np.arange(N, dtype=np.int32), k))
        parts_w.append(np.full(N * k, F32(J), F32))
        d.rates = np.full((1, n_in), F32(reference_input_rate), F32)
    else:
        nu_thr = theta / (J * ce * tau)
        d.ext_rate = float(nu_ext_ratio * nu_thr * ce)
        d.ext_weight = float(F32(J))
    d.pre = np.concatenate(parts_pre)
    d.post = np.concatenate(parts_post)
    d.weight = np.concatenate(parts_w)
    d.delay = np.full(len(d.pre), delay_steps, np.int32)
    d.monitors = list(monitors)
    return d


def random_net(n_neurons=64, n_inputs=32, p=0.2, kind=SYN_VOLTAGE, n_labels=1, seed=1,
               dmin=1, dmax=6, wlo=0.0, whi=2e-3, input_rate=200.0, ext_rate=0.0, ext_w=0.0,
               dt=1e-4, v0=-70e-3, vreset=-70e-3, theta=-55e-3, tref=2e-3, ibg=0.0,
               cm=200e-12, g0=10e-9, n_stimuli=1, grouping=None, monitors=(0, 1),
               scale=None) -> NetDesc:
    """Small random network for parity tests (Bernoulli connectivity, uniform weights and
    delays, optional conductance labels)."""
    rng = np.random.default_rng(seed)
    d = NetDesc(n_neurons=n_neurons, n_inputs=n_inputs, kind=kind, n_labels=n_labels, dt=dt,
                seed=seed, name="random")
    _lif(d, n_neurons, v0, vreset, theta, cm=cm, g0=g0, tref=tref, ibg=ibg)
    U = n_neurons + n_inputs
    mask = rng.random((U, n_neurons)) < p
    pu, post = np.nonzero(mask)
    pre = np.where(pu < n_neurons, pu, -(pu - n_neurons) - 1).astype(np.int32)
    d.pre, d.post = pre, post.astype(np.int32)
    S = len(pre)
    d.weight = rng.uniform(wlo, whi, S).astype(F32)
    if scale is not None:
        d.scale = np.asarray(scale(rng, S), F32) if callable(scale) else np.full(S, F32(scale), F32)
    d.delay = rng.integers(dmin, dmax + 1, S).astype(np.int32)
    if n_labels > 1:
        d.label = rng.integers(0, n_labels, S).astype(np.int32)
    if kind == SYN_CONDUCTANCE:
        taus = [5e-3, 10e-3, 3e-3, 7e-3, 2e-3, 4e-3, 6e-3, 8e-3][:n_labels]
        d.decay_g = np.array([conductance_decay(dt, t) for t in taus], F32)
        d.reversal = np.array([0.0, -80e-3, 0.0, -70e-3, 0.0, -75e-3, 0.0, -90e-3][:n_labels], F32)
    elif kind == SYN_CURRENT:
        # current traces: decay expf(-dt/tau) per label; reversal potentials unused
        taus = [2e-3, 5e-3, 1e-3, 3e-3, 4e-3, 6e-3, 8e-3, 10e-3][:n_labels]
        d.decay_g = np.array([conductance_decay(dt, t) for t in taus], F32)
        d.reversal = np.zeros(n_labels, F32)
    if n_inputs:
        d.rates = rng.uniform(0.5 * input_rate, 1.5 * input_rate, (n_stimuli, n_inputs)).astype(F32)
    d.ext_rate, d.ext_weight = float(ext_rate), float(ext_w)
    d.grouping = grouping
    d.monitors = list(monitors)
    return d


def plain_small(seed=42, dt=1e-4, monitors=(0,)) -> NetDesc:
    """1000 Poisson inputs (20 Hz) -> 1000 LIF, p=0.1 conductance synapses, U(0,1) nS
    weights, delays U{1..10} steps (no plasticity)."""
    rng = np.random.default_rng(seed)
    n, n_in = 1000, 1000
    d = NetDesc(n_neurons=n, n_inputs=n_in, kind=SYN_CONDUCTANCE, n_labels=1, dt=dt, seed=seed,
                name="plain_small")
    _lif(d, n, -70e-3, -70e-3, -50e-3, cm=200e-12, g0=10e-9, tref=2e-3)
    mask = rng.random((n_in, n)) < 0.1
    src, post = np.nonzero(mask)
    d.pre = (-src - 1).astype(np.int32)
    d.post = post.astype(np.int32)
    S = len(d.pre)
    d.weight = rng.uniform(0.0, 1.0, S).astype(F32)
    d.scale = np.full(S, F32(1e-9), F32)
    d.delay = rng.integers(1, 11, S).astype(np.int32)
    d.decay_g = np.array([conductance_decay(dt, 1e-3)], F32)
    d.reversal = np.array([0.0], F32)
    d.rates = np.full((1, n_in), F32(20.0), F32)
    d.monitors = list(monitors)
    return d


def coba_vogels(n_exc=3200, n_inh=800, p=0.02, seed=42, dt=1e-4, delay_steps=8,
                plastic=True, monitors=(0,)) -> NetDesc:
    """Vogels-Abbott COBA network (Examples/VogelsAbbottNet.cpp) with Vogels iSTDP on I->E:
    tau_E 5 ms, tau_I 10 ms, w_E 0.6 nS, w_I 6.7 nS, E_rev 0 / -80 mV."""
    rng = np.random.default_rng(seed)
    N = n_exc + n_inh
    d = NetDesc(n_neurons=N, n_inputs=0, kind=SYN_CONDUCTANCE, n_labels=2, dt=dt, seed=seed,
                name="coba")
    _lif(d, N, -60e-3, -60e-3, -50e-3, cm=200e-12, g0=10e-9, tref=5e-3, ibg=2e-2)
    mask = rng.random((N, N)) < p
    pre, post = np.nonzero(mask)
    d.pre, d.post = pre.astype(np.int32), post.astype(np.int32)
    S = len(pre)
    inh = pre >= n_exc
    d.weight = np.where(inh, F32(0.67), F32(0.06)).astype(F32)
    d.scale = np.full(S, F32(10e-9), F32)
    d.delay = np.full(S, delay_steps, np.int32)
    d.label = inh.astype(np.int32)
    d.decay_g = np.array([conductance_decay(dt, 5e-3), conductance_decay(dt, 10e-3)], F32)
    d.reversal = np.array([0.0, -80e-3], F32)
    if plastic:
        ie = np.nonzero(inh & (post < n_exc))[0].astype(np.int64)
        d.rules.append(Rule(RULE_VOGELS, ie, dict(tau_istdp=20e-3, learningrate=1e-3,
                                                  targetrate=5.0, momentumrate=0.0)))
    d.monitors = list(monitors)
    return d


def large_plastic(n=500_000, out_degree=1000, n_inputs=None, input_out_degree=1000, input_rate=10.0,
                  delay_max=20, J=0.1e-3, J_in=2.0e-3, g=5.0, theta=20e-3, vreset=10e-3, tau=20e-3,
                  tref=2e-3, seed=42, dt=1e-4, plastic=True, monitors=(), generated=True) -> NetDesc:
    """Large plastic network (shrinks with n): 80 % excitatory LIF neurons, each with
    `out_degree` random targets (drawn with replacement), delays U{1..delay_max} steps
    (so G = 1); excitatory weights w ~ U(0, 1) scaled by J, inhibitory ones fixed at 1
    scaled by -g J; n/20 Poisson inputs at `input_rate`, each with `input_out_degree`
    targets, weight 1 scaled by J_in.  Weight-dependent nearest-spike STDP on E->E with
    tau+ 16.8 ms, tau- 33.7 ms, learning rate 0.005, LTD/LTP ratio 1.05 (A- 0.00525 vs
    A+ 0.005), w in [0, 1].

    generated=True: the synapses are drawn on the device by the Philox fixed out-degree
    rule (NetDesc.outdegree; no host arrays).  generated=False: host arrays drawn with
    numpy (another realisation of the same network statistics)."""
    ne = int(round(0.8 * n))
    n_in = n // 20 if n_inputs is None else n_inputs
    d = NetDesc(n_neurons=n, n_inputs=n_in, kind=SYN_VOLTAGE, dt=dt, seed=seed, name="large_plastic")
    _lif(d, n, 0.0, vreset, theta, cm=200e-12, g0=200e-12 / tau, tref=tref)
    d.rates = np.full((1, n_in), F32(input_rate), F32)
    d.monitors = list(monitors)
    stdp = wd_params(a_minus=1.0, a_plus=1.0, tau_minus=33.7e-3, tau_plus=16.8e-3, lam=0.005, alpha=1.05,
                     w_max=1.0, nearest=1)
    if generated:
        d.outdegree = OutDegree(n_exc=ne, out_degree=out_degree, input_out_degree=input_out_degree,
                                delay_max=delay_max, scale_exc=float(F32(J)), scale_inh=float(F32(-g * J)),
                                scale_input=float(F32(J_in)), weight_inh=1.0, weight_input=1.0,
                                ee_stdp=stdp if plastic else None)
        return d
    rng = np.random.default_rng(seed)
    pre = np.repeat(np.arange(n, dtype=np.int32), out_degree)
    post = rng.integers(0, n, size=n * out_degree, dtype=np.int64).astype(np.int32)
    ipre = np.repeat(np.arange(n_in, dtype=np.int32), input_out_degree)
    ipost = rng.integers(0, n, size=n_in * input_out_degree, dtype=np.int64).astype(np.int32)
    exc = pre < ne
    w = np.where(exc, rng.random(len(pre)).astype(F32), F32(1.0)).astype(F32)
    sc = np.where(exc, F32(J), F32(-g * J)).astype(F32)
    d.pre = np.concatenate([pre, (-ipre - 1).astype(np.int32)])
    d.post = np.concatenate([post, ipost])
    d.weight = np.concatenate([w, np.ones(len(ipre), F32)])
    d.scale = np.concatenate([sc, np.full(len(ipre), F32(J_in), F32)])
    d.delay = rng.integers(1, delay_max + 1, size=len(d.pre), dtype=np.int64).astype(np.int32)
    if plastic:
        ee = np.nonzero(exc & (post < ne))[0].astype(np.int64)
        d.rules.append(Rule(RULE_WD, ee, stdp))
    return d


def wd_params(a_minus=1.0, a_plus=1.0, tau_minus=0.02, tau_plus=0.02, lam=1e-2, alpha=2.02,
              w_max=0.3e-3, nearest=0):
    return dict(a_minus=a_minus, a_plus=a_plus, tau_minus=tau_minus, tau_plus=tau_plus,
                **{"lambda": lam}, alpha=alpha, w_max=w_max, nearest_spike_only=nearest)


def evans_params(tau_D=0.005, alpha_D=0.5, alpha_C=0.5, tau_C=0.004, rho=0.1):
    return dict(decay_term_tau_D=tau_D, model_parameter_alpha_D=alpha_D,
                synaptic_neurotransmitter_concentration_alpha_C=alpha_C,
                decay_term_tau_C=tau_C, learning_rate_rho=rho)


def custom_params(a_minus=1.0, a_plus=1.0, a_star=0.0, tau_minus=0.02, tau_plus=0.02, w_max=1.0,
                  learning_rate=1e-3, nearest=0):
    return dict(a_minus=a_minus, a_plus=a_plus, a_star=a_star, tau_minus=tau_minus, tau_plus=tau_plus,
                w_max=w_max, learning_rate=learning_rate, nearest_spike_only=nearest)


def vogels_params(tau=0.02, lr=4e-4, rate=10.0):
    return dict(tau_istdp=tau, learningrate=lr, targetrate=rate, momentumrate=0.0)


__all__ = ["brunel", "random_net", "plain_small", "coba_vogels", "large_plastic", "wd_params", "evans_params", "custom_params",
           "vogels_params", "Rule", "RULE_WD", "RULE_VOGELS", "RULE_EVANS"]
