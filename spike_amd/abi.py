"""ctypes binding of the C ABI in include/spike_hip.h and include/spike_mlp.h (libspike_hip.so).

This module is plumbing for tests and the benchmark: the product's host interface is
Spike's C++ front-end, whose HIP backend calls the same entry points.  Loading fails
loudly when the shared library has not been built; there is no fallback path.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import weakref
from typing import Optional

import numpy as np

from .netdesc import NetDesc, RULE_WD, RULE_VOGELS, RULE_EVANS, RULE_CUSTOM, RULE_WEIGHTNORM

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libspike_hip.so")

SPK_RUN_PLASTICITY = 1
SPK_RUN_FORCE_LAUNCHES = 2


# Handles still open when the interpreter exits are closed by an atexit hook, while the HIP
# runtime is certainly still up (finalizers that run during interpreter teardown might
# otherwise call into it after its own shutdown began).  Batches close before networks.
_live = weakref.WeakValueDictionary()
_live_seq = [0]


def _track(obj, rank: int):
    _live_seq[0] += 1
    _live[(rank, _live_seq[0])] = obj


@atexit.register
def close_all():
    """Close every network, batch and surrogate handle that is still open."""
    for key in sorted(list(_live.keys())):
        obj = _live.get(key)
        if obj is not None:
            try:
                obj.close()
            except Exception:
                pass


class SpikeError(RuntimeError):
    pass


class WDParams(C.Structure):
    _fields_ = [("a_minus", C.c_float), ("a_plus", C.c_float), ("tau_minus", C.c_float),
                ("tau_plus", C.c_float), ("lambda_", C.c_float), ("alpha", C.c_float),
                ("w_max", C.c_float), ("nearest_spike_only", C.c_int)]


class VogelsParams(C.Structure):
    _fields_ = [("tau_istdp", C.c_float), ("learningrate", C.c_float),
                ("targetrate", C.c_float), ("momentumrate", C.c_float)]


class EvansParams(C.Structure):
    _fields_ = [("decay_term_tau_D", C.c_float), ("model_parameter_alpha_D", C.c_float),
                ("synaptic_neurotransmitter_concentration_alpha_C", C.c_float),
                ("decay_term_tau_C", C.c_float), ("learning_rate_rho", C.c_float)]


class CustomParams(C.Structure):
    _fields_ = [("a_minus", C.c_float), ("a_plus", C.c_float), ("a_star", C.c_float),
                ("tau_minus", C.c_float), ("tau_plus", C.c_float), ("w_max", C.c_float),
                ("learning_rate", C.c_float), ("nearest_spike_only", C.c_int)]


class Info(C.Structure):
    _fields_ = [("n_neurons", C.c_int), ("n_inputs", C.c_int), ("n_labels", C.c_int),
                ("synapse_kind", C.c_int), ("timestep_grouping", C.c_int),
                ("ring_slots", C.c_int), ("history_words", C.c_int), ("frac_bits", C.c_int),
                ("min_delay", C.c_int), ("max_delay", C.c_int), ("n_synapses", C.c_int64),
                ("n_rules", C.c_int), ("n_monitors", C.c_int), ("partitions", C.c_int),
                ("fused", C.c_int), ("record_bytes", C.c_int), ("lds_bytes", C.c_int)]


class Timing(C.Structure):
    _fields_ = [("total_ms", C.c_double), ("kernel_ms", C.c_double),
                ("kernel_launches", C.c_int64), ("groups", C.c_int64),
                ("spikes_neurons", C.c_int64), ("spikes_inputs", C.c_int64),
                ("synaptic_events", C.c_int64), ("external_events", C.c_int64)]


P = C.c_void_p
fp = C.POINTER(C.c_float)
ip = C.POINTER(C.c_int32)
lp = C.POINTER(C.c_int64)
up = C.POINTER(C.c_uint32)

# name -> (restype, argtypes)
class BrunelParams(C.Structure):
    """spk_brunel_params of include/spike_hip.h."""
    _fields_ = [("n_exc", C.c_int), ("n_inh", C.c_int), ("ce", C.c_int), ("ci", C.c_int),
                ("delay_steps", C.c_int), ("weight_exc", C.c_float), ("weight_inh", C.c_float),
                ("v0", C.c_float), ("vreset", C.c_float), ("threshold", C.c_float),
                ("decay_c", C.c_float), ("resist", C.c_float), ("refractory_s", C.c_float),
                ("ext_rate_hz", C.c_float), ("ext_weight", C.c_float), ("timestep", C.c_float),
                ("seed", C.c_uint64)]


class OutDegreeParams(C.Structure):
    """spk_outdegree_params of include/spike_hip.h."""
    _fields_ = [("seed", C.c_uint64), ("n_exc", C.c_int), ("out_degree", C.c_int),
                ("input_out_degree", C.c_int), ("delay_max", C.c_int), ("scale_exc", C.c_float),
                ("scale_inh", C.c_float), ("scale_input", C.c_float), ("weight_inh", C.c_float),
                ("weight_input", C.c_float), ("ee_stdp", C.c_int), ("stdp", WDParams)]


class Stats(C.Structure):
    """spk_stats of include/spike_hip.h."""
    _fields_ = [("rate_hz", C.c_double), ("isi_cv", C.c_double), ("synchrony", C.c_double),
                ("mean_bin_count", C.c_double), ("spikes", C.c_int64), ("cv_neurons", C.c_int64),
                ("bins", C.c_int64), ("steps", C.c_int64)]


SIGNATURES = {
    "spk_last_error": (C.c_char_p, []),
    "spk_abi_version": (C.c_int, []),
    "spk_device_count": (C.c_int, []),
    "spk_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.POINTER(P)]),
    "spk_destroy": (C.c_int, [P]),
    "spk_set_stream": (C.c_int, [P, P]),
    "spk_set_lif": (C.c_int, [P, fp, fp, fp, fp, fp, C.c_float, C.c_float]),
    "spk_set_poisson_rates": (C.c_int, [P, fp, C.c_int]),
    "spk_select_stimulus": (C.c_int, [P, C.c_int]),
    "spk_set_generator_stimuli": (C.c_int, [P, ip, C.c_int, ip, fp]),
    "spk_set_generator_onset": (C.c_int, [P, C.c_float]),
    "spk_add_rate_monitor": (C.c_int, [P, C.c_int, C.POINTER(C.c_int)]),
    "spk_rate_monitor_counts": (C.c_int, [P, C.c_int, ip]),
    "spk_rate_monitor_reset": (C.c_int, [P, C.c_int]),
    "spk_set_seed": (C.c_int, [P, C.c_uint64]),
    "spk_set_external_drive": (C.c_int, [P, C.c_float, C.c_float]),
    "spk_set_labels": (C.c_int, [P, fp, fp]),
    "spk_set_synapses": (C.c_int, [P, C.c_int64, ip, ip, fp, fp, ip, ip]),
    "spk_add_wd_stdp": (C.c_int, [P, lp, C.c_int64, C.POINTER(WDParams), C.POINTER(C.c_int)]),
    "spk_add_vogels_stdp": (C.c_int, [P, lp, C.c_int64, C.POINTER(VogelsParams), C.POINTER(C.c_int)]),
    "spk_add_evans_stdp": (C.c_int, [P, lp, C.c_int64, C.POINTER(EvansParams), C.POINTER(C.c_int)]),
    "spk_add_custom_stdp": (C.c_int, [P, lp, C.c_int64, C.POINTER(CustomParams), C.POINTER(C.c_int)]),
    "spk_add_weightnorm": (C.c_int, [P, lp, C.c_int64, fp, C.POINTER(C.c_uint8), C.POINTER(C.c_int)]),
    "spk_weightnorm_set_initial": (C.c_int, [P, C.c_int, fp]),
    "spk_add_monitor": (C.c_int, [P, C.c_int, C.c_int64, C.POINTER(C.c_int)]),
    "spk_prepare": (C.c_int, [P, C.c_int]),
    "spk_set_timestep_grouping": (C.c_int, [P, C.c_int]),
    "spk_set_partitions": (C.c_int, [P, C.c_int]),
    "spk_fused_status": (C.c_char_p, [P]),
    "spk_set_profiling": (C.c_int, [P, C.c_int]),
    "spk_get_phase_profile": (C.c_int, [P, C.POINTER(C.c_double), C.c_int]),
    "spk_get_info": (C.c_int, [P, C.POINTER(Info)]),
    "spk_reset_state": (C.c_int, [P]),
    "spk_reset_inputs": (C.c_int, [P]),
    "spk_update_neurons": (C.c_int, [P, C.c_float]),
    "spk_update_inputs": (C.c_int, [P, C.c_float]),
    "spk_update_plasticity": (C.c_int, [P, C.c_int, C.c_float]),
    "spk_update_synapses": (C.c_int, [P, C.c_float]),
    "spk_collect_spikes": (C.c_int, [P, C.c_int, C.c_float]),
    "spk_run": (C.c_int, [P, C.c_float, C.c_int64, C.c_int]),
    "spk_last_timing": (C.c_int, [P, C.POINTER(Timing)]),
    "spk_step_group": (C.c_int, [P, C.c_float, C.c_int]),
    "spk_get_counters": (C.c_int, [P, C.POINTER(Timing)]),
    "spk_monitor_reset": (C.c_int, [P, C.c_int]),
    "spk_stats_begin": (C.c_int, [P, C.c_int, C.c_int64]),
    "spk_create_brunel": (C.c_int, [C.POINTER(BrunelParams), C.c_int, C.POINTER(P)]),
    "spk_set_outdegree_synapses": (C.c_int, [P, C.POINTER(OutDegreeParams)]),
    "spk_get_rule_size": (C.c_int, [P, C.c_int, C.POINTER(C.c_int64)]),
    "spk_batch_create": (C.c_int, [C.POINTER(P), C.c_int, C.c_int, C.POINTER(P)]),
    "spk_batch_destroy": (C.c_int, [P]),
    "spk_batch_set_cost_model": (C.c_int, [P, C.c_double, C.c_double, C.c_double]),
    "spk_batch_workgroups": (C.c_int, [P]),
    "spk_batch_threads": (C.c_int, [P]),
    "spk_batch_set_costs": (C.c_int, [P, C.POINTER(C.c_double)]),
    "spk_batch_run": (C.c_int, [P, C.POINTER(C.c_float), C.POINTER(C.c_int64), C.c_int]),
    "spk_batch_last_ms": (C.c_int, [P, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "spk_batch_set_graph": (C.c_int, [P, C.c_int]),
    "spk_batch_graph_info": (C.c_int, [P, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "spk_stats_get": (C.c_int, [P, C.POINTER(Stats)]),
    "spk_stats_end": (C.c_int, [P]),
    "spk_synchronize": (C.c_int, [P]),
    "spk_monitor_device_count": (C.c_int, [P, C.c_int, lp]),
    "spk_monitor_drain": (C.c_int, [P, C.c_int, ip, fp, C.c_int64, lp]),
    "spk_monitor_take": (C.c_int, [P, C.c_int, ip, fp, C.c_int64, lp]),
    "spk_get_weights": (C.c_int, [P, fp]),
    "spk_set_weights": (C.c_int, [P, fp]),
    "spk_get_membrane": (C.c_int, [P, fp]),
    "spk_get_last_spike": (C.c_int, [P, C.c_int, fp]),
    "spk_get_conductances": (C.c_int, [P, fp]),
    "spk_get_ring": (C.c_int, [P, lp]),
    "spk_get_history": (C.c_int, [P, C.c_int, up]),
    "spk_get_traces": (C.c_int, [P, C.c_int, fp, fp]),
    # include/spike_mlp.h
    "spk_set_weight_format": (C.c_int, [P, C.c_int]),
    "spk_mlp_param_count": (C.c_int, []),
    "spk_mlp_create": (C.c_int, [C.c_int, C.c_int, C.POINTER(P)]),
    "spk_mlp_destroy": (C.c_int, [P]),
    "spk_mlp_set_params": (C.c_int, [P, fp]),
    "spk_mlp_get_params": (C.c_int, [P, fp]),
    "spk_mlp_reset_optimizer": (C.c_int, [P]),
    "spk_mlp_set_data": (C.c_int, [P, fp, fp, C.c_int64]),
    "spk_mlp_set_loss_weights": (C.c_int, [P, fp]),
    "spk_mlp_gradients": (C.c_int, [P, C.c_int, fp, fp]),
    "spk_mlp_apply_gradients": (C.c_int, [P, fp, C.c_float]),
    "spk_mlp_train": (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.c_uint64, C.c_float, C.c_float, C.c_int,
                                C.c_int, fp]),
    "spk_mlp_predict": (C.c_int, [P, fp, C.c_int64, fp]),
    "spk_mlp_predict_noalloc": (C.c_int, [P, fp, C.c_int64, fp]),
    "spk_mlp_last_timing": (C.c_int, [P, C.c_void_p]),
}

_lib: Optional[C.CDLL] = None


def load(path: str = LIB_PATH) -> C.CDLL:
    """Load libspike_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise SpikeError(f"{path} is missing: build it with __graft_entry__.build()")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _ptr(a: Optional[np.ndarray], ctype):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ctype))


def check(rc: int):
    if rc != 0:
        msg = load().spk_last_error().decode(errors="replace")
        raise SpikeError(f"spike_hip error {rc}: {msg}")


class Network:
    """A network on the GPU built from a NetDesc through the C ABI."""

    def __init__(self, d: NetDesc, monitor_capacity: Optional[int] = None, partitions: int = 0):
        lib = load()
        self.lib = lib
        self.desc = d
        h = P()
        check(lib.spk_create(d.n_neurons, d.n_inputs, d.kind, d.n_labels, float(d.dt), C.byref(h)))
        self.h = h
        _track(self, 2)
        try:
            if partitions:
                check(lib.spk_set_partitions(h, int(partitions)))
            self._build(d, monitor_capacity)
        except Exception:
            lib.spk_destroy(self.h)
            self.h = None
            raise

    def _build(self, d: NetDesc, monitor_capacity):
        lib, h = self.lib, self.h
        f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)
        i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
        self._keep = []
        if d.n_neurons:
            arrs = [f32(x) for x in (d.v0, d.vreset, d.threshold, d.decay_c, d.resist)]
            check(lib.spk_set_lif(h, *[_ptr(a, C.c_float) for a in arrs],
                                  float(d.refractory), float(d.background)))
        if d.n_inputs and d.stimuli is not None:
            counts = i32([len(ids) for ids, _ in d.stimuli])
            ids = i32(np.concatenate([np.asarray(i) for i, _ in d.stimuli]) if len(d.stimuli) else [])
            times = f32(np.concatenate([np.asarray(t) for _, t in d.stimuli]) if len(d.stimuli) else [])
            self._keep += [counts, ids, times]
            check(lib.spk_set_generator_stimuli(h, _ptr(counts, C.c_int32), len(counts), _ptr(ids, C.c_int32),
                                                _ptr(times, C.c_float)))
            check(lib.spk_set_generator_onset(h, float(d.onset)))
        elif d.n_inputs:
            rates = f32(d.rates if d.rates is not None else np.zeros((1, d.n_inputs)))
            rates = rates.reshape(-1, d.n_inputs)
            check(lib.spk_set_poisson_rates(h, _ptr(rates, C.c_float), rates.shape[0]))
        check(lib.spk_set_seed(h, int(d.seed)))
        if d.weight_format:
            check(lib.spk_set_weight_format(h, int(d.weight_format)))
        if d.ext_rate:
            check(lib.spk_set_external_drive(h, float(d.ext_rate), float(d.ext_weight)))
        if d.decay_g is not None:
            dg, rv = f32(d.decay_g), f32(d.reversal)
            check(lib.spk_set_labels(h, _ptr(dg, C.c_float), _ptr(rv, C.c_float)))
        if d.outdegree is not None:
            o = d.outdegree
            p = OutDegreeParams(seed=int(d.seed), n_exc=o.n_exc, out_degree=o.out_degree,
                                input_out_degree=o.input_out_degree, delay_max=o.delay_max,
                                scale_exc=o.scale_exc, scale_inh=o.scale_inh, scale_input=o.scale_input,
                                weight_inh=o.weight_inh, weight_input=o.weight_input,
                                ee_stdp=1 if o.ee_stdp else 0)
            if o.ee_stdp:
                p.stdp = WDParams(**{("lambda_" if k == "lambda" else k): v for k, v in o.ee_stdp.items()})
            check(lib.spk_set_outdegree_synapses(h, C.byref(p)))
        elif d.n_synapses:
            pre, post, dl = i32(d.pre), i32(d.post), i32(d.delay)
            w = f32(d.weight)
            sc = None if d.scale is None else f32(d.scale)
            lb = None if d.label is None else i32(d.label)
            check(lib.spk_set_synapses(h, d.n_synapses, _ptr(pre, C.c_int32), _ptr(post, C.c_int32),
                                       _ptr(w, C.c_float), _ptr(sc, C.c_float), _ptr(dl, C.c_int32),
                                       _ptr(lb, C.c_int32)))
        self.rule_ids = [0] if d.outdegree is not None and d.outdegree.ee_stdp else []
        for r in d.rules:
            syn = np.ascontiguousarray(r.synapses, dtype=np.int64)
            rid = C.c_int()
            if r.kind == RULE_WD:
                p = WDParams(**{("lambda_" if k == "lambda" else k): v for k, v in r.params.items()})
                check(lib.spk_add_wd_stdp(h, _ptr(syn, C.c_int64), len(syn), C.byref(p), C.byref(rid)))
            elif r.kind == RULE_VOGELS:
                p = VogelsParams(**r.params)
                check(lib.spk_add_vogels_stdp(h, _ptr(syn, C.c_int64), len(syn), C.byref(p), C.byref(rid)))
            elif r.kind == RULE_EVANS:
                p = EvansParams(**r.params)
                check(lib.spk_add_evans_stdp(h, _ptr(syn, C.c_int64), len(syn), C.byref(p), C.byref(rid)))
            elif r.kind == RULE_WEIGHTNORM:
                sq = f32(r.params["sum_sq"])
                ins = np.ascontiguousarray(r.params["in_set"], dtype=np.uint8)
                self._keep += [sq, ins]
                check(lib.spk_add_weightnorm(h, _ptr(syn, C.c_int64), len(syn), _ptr(sq, C.c_float),
                                             ins.ctypes.data_as(C.POINTER(C.c_uint8)), C.byref(rid)))
            elif r.kind == RULE_CUSTOM:
                p = CustomParams(**r.params)
                check(lib.spk_add_custom_stdp(h, _ptr(syn, C.c_int64), len(syn), C.byref(p), C.byref(rid)))
            self.rule_ids.append(rid.value)
        self.monitor_ids = []
        for pop in d.monitors:
            npop = d.n_neurons if pop == 0 else d.n_inputs
            cap = monitor_capacity or max(1, 60 * npop)
            mid = C.c_int()
            check(lib.spk_add_monitor(h, int(pop), int(cap), C.byref(mid)))
            self.monitor_ids.append(mid.value)
        self.rate_ids = []
        for pop in getattr(d, "rate_monitors", []):
            mid = C.c_int()
            check(lib.spk_add_rate_monitor(h, int(pop), C.byref(mid)))
            self.rate_ids.append((mid.value, int(pop)))
        check(lib.spk_prepare(h, d.timestep_grouping()))

    def rate_counts(self, k: int = 0) -> np.ndarray:
        """Spike counts per unit of rate monitor k since prepare or its last reset."""
        mid, pop = self.rate_ids[k]
        out = np.zeros(self.desc.n_neurons if pop == 0 else self.desc.n_inputs, np.int32)
        check(self.lib.spk_rate_monitor_counts(self.h, mid, _ptr(out, C.c_int32)))
        return out

    def rate_reset(self, k: int = 0):
        check(self.lib.spk_rate_monitor_reset(self.h, self.rate_ids[k][0]))

    def close(self):
        if getattr(self, "h", None):
            self.lib.spk_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- control -------------------------------------------------------------------
    def info(self) -> Info:
        i = Info()
        check(self.lib.spk_get_info(self.h, C.byref(i)))
        return i

    def reset_state(self):
        check(self.lib.spk_reset_state(self.h))

    def set_profiling(self, on: bool = True):
        check(self.lib.spk_set_profiling(self.h, 1 if on else 0))

    def phase_profile(self):
        out = (C.c_double * 12)()
        check(self.lib.spk_get_phase_profile(self.h, out, 12))
        names = ["inputs_lifinput", "lif", "publish", "arrive_draw", "wait", "plasticity",
                 "dlv_heads", "dlv_scan_stage", "dlv_synapses", "lif_recurrence", "s10", "s11"]
        return dict(zip(names, list(out)))

    def fused_status(self) -> str:
        return self.lib.spk_fused_status(self.h).decode()

    def set_grouping(self, g: int):
        check(self.lib.spk_set_timestep_grouping(self.h, int(g)))

    def run(self, start: float, ngroups: int, plasticity: bool = True, force_launches=False) -> Timing:
        flags = (SPK_RUN_PLASTICITY if plasticity else 0) | (SPK_RUN_FORCE_LAUNCHES if force_launches else 0)
        check(self.lib.spk_run(self.h, float(start), int(ngroups), flags))
        t = Timing()
        check(self.lib.spk_last_timing(self.h, C.byref(t)))
        return t

    def step_group(self, t: float, plasticity: bool = True, force_launches=False):
        """One group enqueued without waiting (the backend's per-group call)."""
        flags = (SPK_RUN_PLASTICITY if plasticity else 0) | (SPK_RUN_FORCE_LAUNCHES if force_launches else 0)
        check(self.lib.spk_step_group(self.h, float(t), flags))

    def stats_begin(self, bin_steps: int = 10, max_steps: int = 1 << 20):
        check(self.lib.spk_stats_begin(self.h, int(bin_steps), int(max_steps)))

    def stats(self) -> Stats:
        st = Stats()
        check(self.lib.spk_stats_get(self.h, C.byref(st)))
        return st

    def stats_end(self):
        check(self.lib.spk_stats_end(self.h))

    def counters(self) -> Timing:
        t = Timing()
        check(self.lib.spk_get_counters(self.h, C.byref(t)))
        return t

    def synchronize(self):
        check(self.lib.spk_synchronize(self.h))

    def step_phases(self, t: float, plasticity: bool = True):
        """One group through the per-phase entry points, in SpikingModel order."""
        lib, h = self.lib, self.h
        check(lib.spk_update_neurons(h, float(t)))
        check(lib.spk_update_inputs(h, float(t)))
        if plasticity:
            for r in self.rule_ids:
                check(lib.spk_update_plasticity(h, r, float(t)))
        check(lib.spk_update_synapses(h, float(t)))
        for m in self.monitor_ids:
            check(lib.spk_collect_spikes(h, m, float(t)))

    # -- read-out ------------------------------------------------------------------
    def membrane(self) -> np.ndarray:
        v = np.empty(self.desc.n_neurons, np.float32)
        check(self.lib.spk_get_membrane(self.h, _ptr(v, C.c_float)))
        return v

    def last_spike(self, pop=0) -> np.ndarray:
        n = self.desc.n_neurons if pop == 0 else self.desc.n_inputs
        v = np.empty(n, np.float32)
        check(self.lib.spk_get_last_spike(self.h, pop, _ptr(v, C.c_float)))
        return v

    def weights(self) -> np.ndarray:
        w = np.empty(self.desc.n_synapses, np.float32)
        check(self.lib.spk_get_weights(self.h, _ptr(w, C.c_float)))
        return w

    def set_weights(self, w: np.ndarray):
        w = np.ascontiguousarray(w, dtype=np.float32)
        check(self.lib.spk_set_weights(self.h, _ptr(w, C.c_float)))

    def conductances(self) -> np.ndarray:
        g = np.empty(self.desc.n_neurons * self.desc.n_labels, np.float32)
        check(self.lib.spk_get_conductances(self.h, _ptr(g, C.c_float)))
        return g

    def ring(self) -> np.ndarray:
        i = self.info()
        r = np.empty(i.ring_slots * self.desc.n_neurons * self.desc.n_labels, np.int64)
        check(self.lib.spk_get_ring(self.h, _ptr(r, C.c_int64)))
        return r

    def history(self, pop=0) -> np.ndarray:
        i = self.info()
        n = self.desc.n_neurons if pop == 0 else self.desc.n_inputs
        r = np.empty(i.history_words * n, np.uint32)
        check(self.lib.spk_get_history(self.h, pop, _ptr(r, C.c_uint32)))
        return r

    def rule_size(self, rule: int) -> int:
        c = C.c_int64()
        check(self.lib.spk_get_rule_size(self.h, int(rule), C.byref(c)))
        return c.value

    def traces(self, rule: int):
        n = self.rule_size(rule)
        a = np.empty(n, np.float32)
        b = np.empty(n, np.float32)
        check(self.lib.spk_get_traces(self.h, rule, _ptr(a, C.c_float), _ptr(b, C.c_float)))
        return a, b

    def take_spikes(self, monitor: int = 0):
        """Records of a monitor: those spk_run drained to the host, then those still in
        the device store (per-phase collection), each sorted by (time, id)."""
        hi, ht = self._take_host(monitor)
        c = C.c_int64()
        check(self.lib.spk_monitor_device_count(self.h, monitor, C.byref(c)))
        if c.value == 0:
            return hi, ht
        di, dt = self.drain_spikes(monitor, cap=int(c.value))
        return np.concatenate([hi, di]), np.concatenate([ht, dt])

    def _take_host(self, monitor: int = 0):
        n = C.c_int64()
        check(self.lib.spk_monitor_take(self.h, monitor, None, None, 0, C.byref(n)))
        ids = np.empty(n.value, np.int32)
        times = np.empty(n.value, np.float32)
        check(self.lib.spk_monitor_take(self.h, monitor, _ptr(ids, C.c_int32), _ptr(times, C.c_float),
                                        n.value, C.byref(n)))
        return ids, times

    def drain_spikes(self, monitor: int = 0, cap: int = 1 << 24):
        ids = np.empty(cap, np.int32)
        times = np.empty(cap, np.float32)
        n = C.c_int64()
        check(self.lib.spk_monitor_drain(self.h, monitor, _ptr(ids, C.c_int32), _ptr(times, C.c_float),
                                         cap, C.byref(n)))
        return ids[: n.value].copy(), times[: n.value].copy()


def generated_brunel(params: "BrunelParams", partitions: int = 0) -> Network:
    """A Brunel network whose connectivity the library generates on the device
    (spk_create_brunel); it runs on the persistent kernel only."""
    lib = load()
    h = P()
    check(lib.spk_create_brunel(C.byref(params), int(partitions), C.byref(h)))
    net = Network.__new__(Network)
    net.lib = lib
    net.h = h
    n = params.n_exc + params.n_inh
    net.desc = NetDesc(n_neurons=n, n_inputs=0, dt=float(params.timestep), grouping=params.delay_steps,
                       name="brunel_generated")
    net.rule_ids = []
    net.monitor_ids = []
    net._keep = []
    _track(net, 2)
    return net


class Batch:
    """Many networks in one persistent launch (spk_batch_*)."""

    def __init__(self, nets, workgroups: int = 0):
        self.lib = load()
        self.nets = list(nets)
        arr = (P * len(self.nets))(*[n.h for n in self.nets])
        h = P()
        check(self.lib.spk_batch_create(arr, len(self.nets), int(workgroups), C.byref(h)))
        self.h = h
        _track(self, 0)

    @property
    def workgroups(self) -> int:
        return self.lib.spk_batch_workgroups(self.h)

    @property
    def threads(self) -> int:
        return self.lib.spk_batch_threads(self.h)

    def set_cost_model(self, fixed_us: float, per_neuron_us: float, per_synapse_us: float):
        check(self.lib.spk_batch_set_cost_model(self.h, fixed_us, per_neuron_us, per_synapse_us))

    def set_costs(self, cost_per_group_us):
        if cost_per_group_us is None:
            check(self.lib.spk_batch_set_costs(self.h, None))
            return
        c = np.ascontiguousarray(cost_per_group_us, dtype=np.float64)
        self._costs = c
        check(self.lib.spk_batch_set_costs(self.h, _ptr(c, C.c_double)))

    def set_graph(self, on: bool = True):
        """Run through a captured hipGraph (default) or launch directly."""
        check(self.lib.spk_batch_set_graph(self.h, 1 if on else 0))

    def graph_info(self):
        """(captures, replays) of the batch's hipGraph."""
        c, r = C.c_int64(), C.c_int64()
        check(self.lib.spk_batch_graph_info(self.h, C.byref(c), C.byref(r)))
        return c.value, r.value

    def run(self, starts, ngroups):
        st = np.ascontiguousarray(starts, dtype=np.float32)
        ng = np.ascontiguousarray(ngroups, dtype=np.int64)
        check(self.lib.spk_batch_run(self.h, _ptr(st, C.c_float), _ptr(ng, C.c_int64), 0))
        km, em = C.c_double(), C.c_double()
        check(self.lib.spk_batch_last_ms(self.h, C.byref(km), C.byref(em)))
        return km.value, em.value

    def close(self):
        if getattr(self, "h", None):
            self.lib.spk_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
