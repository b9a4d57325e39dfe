"""ctypes binding of the CPU oracle (oracle/_build/libspike_oracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and bench.py's CPU
baseline as the checker.  Mirrors spike_amd.abi.Network so a parity test can build the
same NetDesc on both.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# SPIKE_ORACLE_LIB selects another build of the same source (the mutation test loads
# oracles with planted mistakes this way)
LIB_PATH = os.environ.get("SPIKE_ORACLE_LIB") or os.path.join(HERE, "_build", "libspike_oracle.so")

_lib = None

P = C.c_void_p
fp = C.POINTER(C.c_float)
dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int32)
lp = C.POINTER(C.c_int64)
up = C.POINTER(C.c_uint32)

SIG = {
    "or_last_error": (C.c_char_p, []),
    "or_create": (P, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int]),
    "or_destroy": (C.c_int, [P]),
    "or_set_threads": (C.c_int, [P, C.c_int]),
    "or_set_lif": (C.c_int, [P, fp, fp, fp, fp, fp, C.c_float, C.c_float]),
    "or_set_poisson_rates": (C.c_int, [P, fp, C.c_int]),
    "or_select_stimulus": (C.c_int, [P, C.c_int]),
    "or_set_generator_stimuli": (C.c_int, [P, ip, C.c_int, ip, fp]),
    "or_set_generator_onset": (C.c_int, [P, C.c_float]),
    "or_set_seed": (C.c_int, [P, C.c_uint64]),
    "or_set_external_drive": (C.c_int, [P, C.c_float, C.c_float]),
    "or_set_weight_format": (C.c_int, [P, C.c_int]),
    "or_round16_bits": (C.c_uint16, [C.c_float, C.c_int]),
    "or_decode16": (C.c_float, [C.c_uint16, C.c_int]),
    "or_set_labels": (C.c_int, [P, fp, fp]),
    "or_set_synapses": (C.c_int, [P, C.c_int64, ip, ip, fp, fp, ip, ip]),
    "or_add_wd_stdp": (C.c_int, [P, lp, C.c_int64, fp, C.c_int]),
    "or_add_vogels_stdp": (C.c_int, [P, lp, C.c_int64, fp]),
    "or_add_evans_stdp": (C.c_int, [P, lp, C.c_int64, fp]),
    "or_add_custom_stdp": (C.c_int, [P, lp, C.c_int64, fp, C.c_int]),
    "or_add_weightnorm": (C.c_int, [P, lp, C.c_int64, fp, C.POINTER(C.c_uint8)]),
    "or_weightnorm_set_initial": (C.c_int, [P, C.c_int, fp]),
    "or_add_monitor": (C.c_int, [P, C.c_int]),
    "or_prepare": (C.c_int, [P, C.c_int]),
    "or_set_timestep_grouping": (C.c_int, [P, C.c_int]),
    "or_get_info": (C.c_int, [P, ip]),
    "or_reset_state": (C.c_int, [P]),
    "or_reset_inputs": (C.c_int, [P]),
    "or_update_neurons": (C.c_int, [P, C.c_float]),
    "or_update_inputs": (C.c_int, [P, C.c_float]),
    "or_update_plasticity": (C.c_int, [P, C.c_int, C.c_float]),
    "or_update_synapses": (C.c_int, [P, C.c_float]),
    "or_collect_spikes": (C.c_int, [P, C.c_int, C.c_float]),
    "or_run": (C.c_int, [P, C.c_float, C.c_int64, C.c_int]),
    "or_get_counters": (C.c_int, [P, lp]),
    "or_monitor_count": (C.c_int64, [P, C.c_int]),
    "or_monitor_take": (C.c_int, [P, C.c_int, ip, fp]),
    "or_get_weights": (C.c_int, [P, dp]),
    "or_set_weights": (C.c_int, [P, dp]),
    "or_get_membrane": (C.c_int, [P, dp]),
    "or_get_last_spike": (C.c_int, [P, C.c_int, fp]),
    "or_get_conductances": (C.c_int, [P, dp]),
    "or_get_ring_fixed": (C.c_int, [P, lp]),
    "or_get_ring_real": (C.c_int, [P, dp]),
    "or_get_history": (C.c_int, [P, C.c_int, up]),
    "or_get_traces": (C.c_int, [P, C.c_int, dp, dp]),
    "or_philox4x32_10": (None, [up, up, up]),
    "or_brunel_connectivity": (C.c_int, [C.c_uint64, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                         C.POINTER(C.c_int32)]),
    "or_outdegree_connectivity": (C.c_int, [C.c_uint64, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                            C.c_float, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int,
                                            C.POINTER(C.c_int32), C.POINTER(C.c_int32), fp, fp,
                                            C.POINTER(C.c_int32)]),
    "or_bernoulli_threshold": (C.c_uint32, [C.c_float, C.c_float]),
    "or_poisson_count_table": (C.c_int, [C.c_double, up, C.c_int]),
    "or_choose_frac_bits": (C.c_int, [C.c_double]),
}


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    for k, (r, a) in SIG.items():
        f = getattr(lib, k)
        f.restype = r
        f.argtypes = a
    _lib = lib
    return lib


def _p(a, t):
    return None if a is None else a.ctypes.data_as(C.POINTER(t))


def check(rc):
    if rc != 0:
        raise RuntimeError("oracle: " + load().or_last_error().decode())


def brunel_connectivity(seed, ne, ni, ce, ci):
    """Presynaptic ids [N, ce+ci] of the Philox fixed in-degree rule."""
    lib = load()
    n = ne + ni
    pre = np.empty((n, ce + ci), np.int32)
    check(lib.or_brunel_connectivity(int(seed), ne, ni, ce, ci, n, _p(pre, C.c_int32)))
    return pre


def outdegree_connectivity(seed, n, n_in, o, threads=8):
    """Synapse arrays (pre, post, weight, scale, delay) of the Philox fixed out-degree
    rule with the parameters of a NetDesc `outdegree` spec."""
    lib = load()
    S = o.n_synapses(n, n_in)
    pre = np.empty(S, np.int32)
    post = np.empty(S, np.int32)
    w = np.empty(S, np.float32)
    sc = np.empty(S, np.float32)
    dl = np.empty(S, np.int32)
    check(lib.or_outdegree_connectivity(int(seed), n, n_in, o.n_exc, o.out_degree, o.input_out_degree,
                                        o.delay_max, o.scale_exc, o.scale_inh, o.scale_input, o.weight_inh,
                                        o.weight_input, int(threads), _p(pre, C.c_int32), _p(post, C.c_int32),
                                        _p(w, C.c_float), _p(sc, C.c_float), _p(dl, C.c_int32)))
    return pre, post, w, sc, dl


def expand_outdegree(d, threads=8):
    """The same network as a NetDesc with host synapse arrays (and its E->E rule listed in
    synapse order): what the library generates on the device for `d.outdegree`."""
    import copy
    from spike_amd.netdesc import Rule, RULE_WD
    o = d.outdegree
    e = copy.copy(d)
    e.outdegree = None
    e.pre, e.post, e.weight, e.scale, e.delay = outdegree_connectivity(d.seed, d.n_neurons, d.n_inputs, o, threads)
    e.label = None
    e.grouping = d.timestep_grouping()
    e.rules = []
    if o.ee_stdp:
        ee = np.flatnonzero((e.pre >= 0) & (e.pre < o.n_exc) & (e.post < o.n_exc)).astype(np.int64)
        e.rules = [Rule(RULE_WD, ee, dict(o.ee_stdp))]
    return e


def philox(ctr, key):
    lib = load()
    c = np.ascontiguousarray(ctr, np.uint32)
    k = np.ascontiguousarray(key, np.uint32)
    o = np.zeros(4, np.uint32)
    lib.or_philox4x32_10(_p(c, C.c_uint32), _p(k, C.c_uint32), _p(o, C.c_uint32))
    return o


class OracleNetwork:
    """The same NetDesc on the CPU oracle (precision 32 or 64)."""

    def __init__(self, d, precision: int = 32, threads: int = 1):
        lib = load()
        self.lib, self.desc, self.precision = lib, d, precision
        h = lib.or_create(d.n_neurons, d.n_inputs, d.kind, d.n_labels, float(d.dt), precision)
        if not h:
            check(1)
        self.h = h
        check(lib.or_set_threads(h, int(threads)))
        f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)
        i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
        if d.n_neurons:
            arrs = [f32(x) for x in (d.v0, d.vreset, d.threshold, d.decay_c, d.resist)]
            check(lib.or_set_lif(h, *[_p(a, C.c_float) for a in arrs], float(d.refractory),
                                 float(d.background)))
        if d.n_inputs and getattr(d, "stimuli", None) is not None:
            counts = i32([len(ids) for ids, _ in d.stimuli])
            ids = i32(np.concatenate([np.asarray(i) for i, _ in d.stimuli]) if len(d.stimuli) else [])
            times = f32(np.concatenate([np.asarray(t) for _, t in d.stimuli]) if len(d.stimuli) else [])
            check(lib.or_set_generator_stimuli(h, _p(counts, C.c_int32), len(counts), _p(ids, C.c_int32),
                                               _p(times, C.c_float)))
            check(lib.or_set_generator_onset(h, float(d.onset)))
        elif d.n_inputs:
            rates = f32(d.rates if d.rates is not None else np.zeros((1, d.n_inputs))).reshape(-1, d.n_inputs)
            check(lib.or_set_poisson_rates(h, _p(rates, C.c_float), rates.shape[0]))
        check(lib.or_set_seed(h, int(d.seed)))
        if getattr(d, "weight_format", 0):
            check(lib.or_set_weight_format(h, int(d.weight_format)))
        if d.ext_rate:
            check(lib.or_set_external_drive(h, float(d.ext_rate), float(d.ext_weight)))
        if d.decay_g is not None:
            dg, rv = f32(d.decay_g), f32(d.reversal)
            check(lib.or_set_labels(h, _p(dg, C.c_float), _p(rv, C.c_float)))
        if d.n_synapses:
            pre, post, dl = i32(d.pre), i32(d.post), i32(d.delay)
            w = f32(d.weight)
            sc = None if d.scale is None else f32(d.scale)
            lb = None if d.label is None else i32(d.label)
            check(lib.or_set_synapses(h, d.n_synapses, _p(pre, C.c_int32), _p(post, C.c_int32),
                                      _p(w, C.c_float), _p(sc, C.c_float), _p(dl, C.c_int32),
                                      _p(lb, C.c_int32)))
        for r in d.rules:
            syn = np.ascontiguousarray(r.synapses, dtype=np.int64)
            pr = r.params
            if r.kind == 0:
                par = f32([pr["a_minus"], pr["a_plus"], pr["tau_minus"], pr["tau_plus"],
                           pr["lambda"], pr["alpha"], pr["w_max"]])
                check(lib.or_add_wd_stdp(h, _p(syn, C.c_int64), len(syn), _p(par, C.c_float),
                                         int(pr.get("nearest_spike_only", 0))))
            elif r.kind == 1:
                par = f32([pr["tau_istdp"], pr["learningrate"], pr["targetrate"], pr.get("momentumrate", 0.0)])
                check(lib.or_add_vogels_stdp(h, _p(syn, C.c_int64), len(syn), _p(par, C.c_float)))
            elif r.kind == 4:
                sq = f32(pr["sum_sq"])
                ins = np.ascontiguousarray(pr["in_set"], dtype=np.uint8)
                check(lib.or_add_weightnorm(h, _p(syn, C.c_int64), len(syn), _p(sq, C.c_float),
                                            ins.ctypes.data_as(C.POINTER(C.c_uint8))))
            elif r.kind == 3:
                par = f32([pr["a_minus"], pr["a_plus"], pr["a_star"], pr["tau_minus"], pr["tau_plus"],
                           pr["w_max"], pr["learning_rate"]])
                check(lib.or_add_custom_stdp(h, _p(syn, C.c_int64), len(syn), _p(par, C.c_float),
                                             int(pr.get("nearest_spike_only", 0))))
            else:
                par = f32([pr["decay_term_tau_D"], pr["model_parameter_alpha_D"],
                           pr["synaptic_neurotransmitter_concentration_alpha_C"],
                           pr["decay_term_tau_C"], pr.get("learning_rate_rho", 0.1)])
                check(lib.or_add_evans_stdp(h, _p(syn, C.c_int64), len(syn), _p(par, C.c_float)))
        for pop in d.monitors:
            lib.or_add_monitor(h, int(pop))
        check(lib.or_prepare(h, d.timestep_grouping()))

    def close(self):
        if getattr(self, "h", None):
            self.lib.or_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        a = np.zeros(8, np.int32)
        check(self.lib.or_get_info(self.h, _p(a, C.c_int32)))
        return dict(G=int(a[0]), B=int(a[1]), W=int(a[2]), frac=int(a[3]), dmin=int(a[4]),
                    dmax=int(a[5]), ext_k=int(a[6]))

    def reset_state(self):
        check(self.lib.or_reset_state(self.h))

    def set_threads(self, threads: int):
        check(self.lib.or_set_threads(self.h, int(threads)))

    def set_grouping(self, g):
        check(self.lib.or_set_timestep_grouping(self.h, int(g)))

    def run(self, start, ngroups, plasticity=True):
        check(self.lib.or_run(self.h, float(start), int(ngroups), 1 if plasticity else 0))

    def step_phases(self, t, plasticity=True):
        lib, h = self.lib, self.h
        check(lib.or_update_neurons(h, float(t)))
        check(lib.or_update_inputs(h, float(t)))
        if plasticity:
            for r in range(len(self.desc.rules)):
                check(lib.or_update_plasticity(h, r, float(t)))
        check(lib.or_update_synapses(h, float(t)))
        for m in range(len(self.desc.monitors)):
            check(lib.or_collect_spikes(h, m, float(t)))

    def counters(self):
        c = np.zeros(4, np.int64)
        check(self.lib.or_get_counters(self.h, _p(c, C.c_int64)))
        return dict(spikes_neurons=int(c[0]), spikes_inputs=int(c[1]), synaptic_events=int(c[2]),
                    external_events=int(c[3]))

    def membrane(self):
        v = np.empty(self.desc.n_neurons, np.float64)
        check(self.lib.or_get_membrane(self.h, _p(v, C.c_double)))
        return v

    def last_spike(self, pop=0):
        n = self.desc.n_neurons if pop == 0 else self.desc.n_inputs
        v = np.empty(n, np.float32)
        check(self.lib.or_get_last_spike(self.h, pop, _p(v, C.c_float)))
        return v

    def weights(self):
        w = np.empty(self.desc.n_synapses, np.float64)
        check(self.lib.or_get_weights(self.h, _p(w, C.c_double)))
        return w

    def set_weights(self, w):
        w = np.ascontiguousarray(w, np.float64)
        assert w.shape == (self.desc.n_synapses,)
        check(self.lib.or_set_weights(self.h, _p(w, C.c_double)))

    def conductances(self):
        g = np.empty(self.desc.n_neurons * self.desc.n_labels, np.float64)
        check(self.lib.or_get_conductances(self.h, _p(g, C.c_double)))
        return g

    def ring(self):
        i = self.info()
        n = i["B"] * self.desc.n_neurons * self.desc.n_labels
        if self.precision == 32:
            r = np.empty(n, np.int64)
            check(self.lib.or_get_ring_fixed(self.h, _p(r, C.c_int64)))
        else:
            r = np.empty(n, np.float64)
            check(self.lib.or_get_ring_real(self.h, _p(r, C.c_double)))
        return r

    def history(self, pop=0):
        i = self.info()
        n = self.desc.n_neurons if pop == 0 else self.desc.n_inputs
        r = np.empty(i["W"] * n, np.uint32)
        check(self.lib.or_get_history(self.h, pop, _p(r, C.c_uint32)))
        return r

    def traces(self, rule):
        n = len(self.desc.rules[rule].synapses)
        a = np.empty(n, np.float64)
        b = np.empty(n, np.float64)
        check(self.lib.or_get_traces(self.h, rule, _p(a, C.c_double), _p(b, C.c_double)))
        return a, b

    def take_spikes(self, monitor=0):
        n = self.lib.or_monitor_count(self.h, monitor)
        ids = np.empty(n, np.int32)
        t = np.empty(n, np.float32)
        check(self.lib.or_monitor_take(self.h, monitor, _p(ids, C.c_int32), _p(t, C.c_float)))
        return ids, t
