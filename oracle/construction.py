"""CPU restatement of Spike's network construction (test infrastructure, not product).

Spike builds connectivity in its C++ front-end with glibc ``rand()`` after ``srand(seed)``
(Synapses.cpp:15-25).  This module restates that construction for small networks so the
tests can check it, draw for draw, against the fixtures the reference front-end itself
wrote (tests/golden/*, made by tests/golden/gen_connectivity.cpp):

  * glibc ``random()`` (TYPE_3 additive feedback generator, the algorithm glibc documents
    for ``srand``/``rand``; third-party, system libc of this image);
  * ``connect_with_sparsity`` (Examples/UtilityFunctions.hpp:51-82);
  * ``Synapses::AddGroup`` for ALL_TO_ALL / ONE_TO_ONE / RANDOM / PAIRWISE connectivity and
    the weight draws (Synapses.cpp:102-388);
  * ``SpikingSynapses::AddGroup`` delays and labels (SpikingSynapses.cpp:62-124) and the
    conductance label assignment (ConductanceSpikingSynapses.cpp:16-60);
  * ``Synapses::sort_synapses`` / ``SpikingSynapses::prepare_backend_early`` (stable bucket
    sort by presynaptic id, inputs reversed first; efferent starts, Synapses.cpp:33-81,
    SpikingSynapses.cpp:12-30) and the plastic index remap (STDPPlasticity.cpp:13-17).

All float arithmetic is done in float32 exactly as the C++ does it (x86-64 SSE, no FMA).
Pure Python: meant for networks of up to ~1e5 draws.
"""
import math

import numpy as np

F = np.float32
RAND_MAX = 2147483647


class GlibcRand:
    """glibc srand()/rand(): r[i] = r[i-3] + r[i-31] (mod 2^32) after a Lehmer seeding of
    r[0..30] (16807 * x mod 2^31-1 by Schrage's method) and 310 discarded outputs; each
    output is r >> 1."""

    def __init__(self, seed):
        seed &= 0xFFFFFFFF
        if seed == 0:
            seed = 1
        r = [0] * 34
        r[0] = seed if seed < 2**31 else seed - 2**32
        for i in range(1, 31):
            hi, lo = divmod(r[i - 1], 127773) if r[i - 1] >= 0 else (-((-r[i - 1]) // 127773), -((-r[i - 1]) % 127773))
            word = 16807 * lo - 2836 * hi
            if word < 0:
                word += 2147483647
            r[i] = word
        for i in range(31, 34):
            r[i] = r[i - 31]
        self.r = [x & 0xFFFFFFFF for x in r]
        for _ in range(310):
            self._next()

    def _next(self):
        r = self.r
        v = (r[-31] + r[-3]) & 0xFFFFFFFF
        r.append(v)
        if len(r) > 64:
            del r[:-34]
        return v >> 1

    def rand(self):
        return self._next()

    def unit(self):
        """(float)rand() / (RAND_MAX), in float32 (Synapses.cpp:214, :322)."""
        return F(F(self.rand()) / F(RAND_MAX))


class Builder:
    """Front-end state of one network under construction (host arrays of Synapses.hpp,
    SpikingSynapses.hpp, Neurons.hpp)."""

    def __init__(self, seed, timestep=0.0001, conductance=False):
        self.rng = GlibcRand(seed)
        self.dt = F(timestep)
        self.conductance = conductance
        self.neuron_groups = []   # (start, n)
        self.input_groups = []
        self.N = 0
        self.Nin = 0
        self.pre, self.post, self.w, self.scale, self.delay, self.label = [], [], [], [], [], []
        self.min_delay = 10 ** 6
        self.max_delay = 0
        self.labels = []          # (Vhat, tau_g) per label
        self.rules = []           # lists of synapse indices (creation order)
        self.rule_kinds = []

    def add_neurons(self, n):
        self.neuron_groups.append((self.N, n))
        self.N += n
        return len(self.neuron_groups) - 1

    def add_inputs(self, n):
        self.input_groups.append((self.Nin, n))
        self.Nin += n
        return -len(self.input_groups)     # CORRECTED_PRESYNAPTIC_ID (Neurons.hpp:34-35)

    def add_rule(self, kind="wd"):
        self.rules.append([])
        self.rule_kinds.append(kind)
        return len(self.rules) - 1

    # Synapses::AddGroup (Synapses.cpp:102-388) + SpikingSynapses::AddGroup (:62-124)
    def add_group(self, pre_group, post_group, kind="all", weight_range=(0.0, 0.0), scale=1.0,
                  delay_range=(0.0001, 0.0001), p=0.0, pairs=None, rules=(), vhat=0.0, tau_g=0.001):
        is_input = pre_group < 0
        if is_input:
            g = -pre_group - 1
            start, n = self.input_groups[g]
            prestart = start if pre_group < -1 else 0   # Synapses.cpp:135-137
            preend = start + n
        else:
            start, n = self.neuron_groups[pre_group]
            prestart = start if pre_group > 0 else 0    # Synapses.cpp:144-146
            preend = start + n
        poststart, pn = self.neuron_groups[post_group]
        postend = poststart + pn

        def enc(i):
            return -i - 1 if is_input else i

        s0 = len(self.pre)
        if kind == "all":
            for i in range(prestart, preend):
                for j in range(poststart, postend):
                    self.pre.append(enc(i))
                    self.post.append(j)
        elif kind == "one":
            for i in range(preend - prestart):
                self.pre.append(enc(prestart + i))
                self.post.append(poststart + i)
        elif kind == "random":
            pf = F(p)
            for i in range(prestart, preend):
                for j in range(poststart, postend):
                    if self.rng.unit() < pf:
                        self.pre.append(enc(i))
                        self.post.append(j)
        elif kind == "pairwise":
            pre_v, post_v = pairs
            for a, b in zip(pre_v, post_v):
                self.pre.append(enc(prestart + int(a)))
                self.post.append(poststart + int(b))
        else:
            raise ValueError(kind)
        s1 = len(self.pre)
        lo, hi = F(weight_range[0]), F(weight_range[1])
        for _ in range(s0, s1):
            self.scale.append(F(scale))
            wv = lo
            if hi != lo:
                wv = F(lo + F(F(hi - lo) * self.rng.unit()))
            self.w.append(wv)
        for r in rules:
            self.rules[r].extend(range(s0, s1))
        # delays (SpikingSynapses.cpp:81-118): int(round(range/timestep)), uniform draw
        d0 = int(round(float(F(F(delay_range[0]) / self.dt))))
        d1 = int(round(float(F(F(delay_range[1]) / self.dt))))
        if d0 < 1 or d1 < 1:
            raise ValueError("Delay range must be at least one timestep.")
        for _ in range(s0, s1):
            dv = F(d0)
            if d0 != d1:
                dv = F(F(d0) + F(F(d1 - d0) * self.rng.unit()))
            d = int(math.floor(float(dv) + 0.5))    # round() of a non-negative value
            self.delay.append(d)
            self.label.append(0)
            self.max_delay = max(self.max_delay, d)
            self.min_delay = min(self.min_delay, d)
        # conductance labels (ConductanceSpikingSynapses.cpp:38-58)
        if self.conductance:
            key = (F(vhat), F(tau_g))
            if not self.labels:
                self.labels.append(key)
            else:
                if key in self.labels:
                    lab = self.labels.index(key)
                else:
                    self.labels.append(key)
                    lab = len(self.labels) - 1
                for s in range(s0, s1):
                    self.label[s] = lab

    def connect_with_sparsity(self, pre_group, post_group, n_pre, n_post, sparseness, **kw):
        """Examples/UtilityFunctions.hpp:51-82: fixed in-degree, sources rand() % n_pre."""
        per_post = int(F(F(sparseness) * F(n_pre)))
        pre_v, post_v = [], []
        for o in range(n_post):
            for _ in range(per_post):
                post_v.append(o)
                pre_v.append(self.rng.rand() % n_pre)
        self.add_group(pre_group, post_group, kind="pairwise", pairs=(pre_v, post_v), **kw)

    # Synapses::sort_synapses + SpikingSynapses::prepare_backend_early
    def finalise(self):
        S = len(self.pre)
        buckets = [[] for _ in range(self.Nin + self.N)]
        for s in range(S):
            buckets[self.Nin + self.pre[s]].append(s)
        sort_idx = np.array([s for b in buckets for s in b], dtype=np.int32)
        rev = np.empty(S, dtype=np.int32)
        rev[sort_idx] = np.arange(S, dtype=np.int32)
        pre = np.array(self.pre, dtype=np.int32)[sort_idx]
        out = {
            "sort_idx": sort_idx,
            "reversesort_idx": rev,
            "pre": pre,
            "post": np.array(self.post, dtype=np.int32)[sort_idx],
            "w": np.array(self.w, dtype=np.float32)[sort_idx],
            "scale": np.array(self.scale, dtype=np.float32)[sort_idx],
            "delay": np.array(self.delay, dtype=np.int32)[sort_idx],
            "label": np.array(self.label, dtype=np.int32)[sort_idx],
            "timestep_grouping": self.min_delay,
            "min_delay": self.min_delay,
            "max_delay": self.max_delay,
            "num_syn_labels": max(1, len(self.labels)),
        }
        n_start = np.zeros(self.N, dtype=np.int32)
        n_count = np.zeros(self.N, dtype=np.int32)
        i_start = np.zeros(self.Nin, dtype=np.int32)
        i_count = np.zeros(self.Nin, dtype=np.int32)
        for s in range(S):
            p = int(pre[s])
            if p < 0:
                i_count[-p - 1] += 1
            else:
                n_count[p] += 1
            if s == 0 or pre[s - 1] != pre[s]:
                if p < 0:
                    i_start[-p - 1] = s
                else:
                    n_start[p] = s
        out.update(neuron_eff_start=n_start, neuron_eff_count=n_count,
                   input_eff_start=i_start, input_eff_count=i_count)
        # STDPPlasticity::prepare_backend_early maps the indices to sorted positions
        # (STDPPlasticity.cpp:13-17); EvansSTDPPlasticity overrides it without calling it
        # (EvansSTDPPlasticity.cpp:19-27), so an Evans rule keeps creation-order indices.
        out["plastic"] = []
        for r, kind in zip(self.rules, self.rule_kinds):
            idx = np.array(r, dtype=np.int64)
            if kind != "evans" and len(idx):
                idx = rev[idx]
            out["plastic"].append(idx.astype(np.int32))
        return out
