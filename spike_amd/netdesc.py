"""Network descriptions: the plain arrays a Spike model hands to its backend.

A `NetDesc` holds exactly what the reference's backend `prepare()` calls read from the
front-end objects (neuron parameter arrays, the synapse arrays, plasticity index sets,
monitors), already converted the way the reference converts them:

* LIF: ``decay_c = dt / (Cm/g0)`` and ``resist = decay_c * (1/g0)`` in float32
  (Spike/Neurons/LIFSpikingNeurons.cpp:29-30, CUDA LIFSpikingNeurons.cu:31-45);
* conductance labels: ``decay_g = expf(-dt/tau_g)`` (CUDA ConductanceSpikingSynapses.cu:76-78);
* delays in timesteps, ``round(U[d0,d1]/dt)`` (Spike/Synapses/SpikingSynapses.cpp:83-118);
* presynaptic ids with inputs encoded as ``-i-1`` (Spike/Neurons/Neurons.hpp:34-35).

The same description feeds the HIP backend (`spike_amd.abi.Network`) and the CPU oracle,
so a parity test builds a network once and runs it twice.
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional

import numpy as np

F32 = np.float32

WEIGHTS_FP32, WEIGHTS_BF16, WEIGHTS_FP16 = 0, 1, 2
SYN_VOLTAGE = 0
SYN_CONDUCTANCE = 1
SYN_CURRENT = 2
POP_NEURONS = 0
POP_INPUTS = 1

RULE_WD = 0
RULE_VOGELS = 1
RULE_EVANS = 2
RULE_CUSTOM = 3
RULE_WEIGHTNORM = 4   # params: sum_sq [n_neurons] float32, in_set [n_neurons] bool


@dataclasses.dataclass
class Rule:
    kind: int
    synapses: np.ndarray          # int64 indices into the synapse arrays
    params: dict


@dataclasses.dataclass
class OutDegree:
    """Fixed out-degree synapses generated on the device (spk_set_outdegree_synapses):
    every neuron has `out_degree` targets and every input `input_out_degree`, drawn with
    replacement from Philox (stream 5), delays U{1..delay_max} steps, excitatory weights
    U[0, 1) in float32 (sources below n_exc), the others weight_inh / weight_input; with
    `ee_stdp` (weight-dependent STDP parameters) the E->E synapses form rule 0."""
    n_exc: int
    out_degree: int
    input_out_degree: int
    delay_max: int
    scale_exc: float
    scale_inh: float
    scale_input: float
    weight_inh: float = 1.0
    weight_input: float = 1.0
    ee_stdp: Optional[dict] = None

    def n_synapses(self, n_neurons: int, n_inputs: int) -> int:
        return n_neurons * self.out_degree + n_inputs * self.input_out_degree


@dataclasses.dataclass
class NetDesc:
    n_neurons: int
    n_inputs: int
    kind: int = SYN_VOLTAGE
    n_labels: int = 1
    dt: float = 1e-4
    # LIF (float32 arrays of n_neurons)
    v0: Optional[np.ndarray] = None
    vreset: Optional[np.ndarray] = None
    threshold: Optional[np.ndarray] = None
    decay_c: Optional[np.ndarray] = None
    resist: Optional[np.ndarray] = None
    refractory: float = 0.002
    background: float = 0.0
    # Poisson inputs: rates[n_stimuli, n_inputs] in Hz
    rates: Optional[np.ndarray] = None
    # or generator inputs: per stimulus (ids int32, times float32 in s); onset adjustment
    stimuli: Optional[List] = None
    onset: float = 0.0
    seed: int = 42
    ext_rate: float = 0.0
    ext_weight: float = 0.0
    # conductance labels
    decay_g: Optional[np.ndarray] = None
    reversal: Optional[np.ndarray] = None
    # synapses
    pre: Optional[np.ndarray] = None
    post: Optional[np.ndarray] = None
    weight: Optional[np.ndarray] = None
    scale: Optional[np.ndarray] = None
    delay: Optional[np.ndarray] = None
    label: Optional[np.ndarray] = None
    rules: List[Rule] = dataclasses.field(default_factory=list)
    monitors: List[int] = dataclasses.field(default_factory=list)
    rate_monitors: List[int] = dataclasses.field(default_factory=list)   # populations
    grouping: Optional[int] = None   # default: minimum delay (SpikingModel.cpp:98)
    name: str = "net"
    weight_format: int = 0        # 0 FP32, 1 BF16, 2 FP16 storage of weights and traces
    outdegree: Optional[OutDegree] = None   # synapses generated on the device instead of pre..label

    @property
    def n_synapses(self) -> int:
        if self.outdegree is not None:
            return self.outdegree.n_synapses(self.n_neurons, self.n_inputs)
        return 0 if self.pre is None else int(len(self.pre))

    def timestep_grouping(self) -> int:
        if self.grouping is not None:
            return int(self.grouping)
        if self.outdegree is not None:
            return 1   # generated delays start at one step
        if self.n_synapses == 0:
            return 1
        return int(np.min(self.delay))

    def groups_for(self, seconds: float) -> int:
        """Group count of SpikingModel::run (SpikingModel.cpp:221-222): the int/int
        division truncates, so a 1 s Brunel run at G=15 covers 9990 steps."""
        steps = int(math.ceil(F32(seconds) / F32(self.dt)))
        return steps // self.timestep_grouping()

    def group_time(self, start: float, s: int) -> np.float32:
        """t = start + s*G*dt in float32 (SpikingModel.cpp:226)."""
        return F32(F32(start) + F32(F32(s * self.timestep_grouping()) * F32(self.dt)))


def lif_constants(n: int, cm: float, g0: float, dt: float):
    """(decay_c, resist) of a LIF group as the reference's host code computes them."""
    tau = F32(F32(cm) / F32(g0))
    r = F32(F32(1.0) / F32(g0))
    c = F32(F32(dt) / tau)
    rp = F32(c * r)
    return np.full(n, c, dtype=F32), np.full(n, rp, dtype=F32)


def conductance_decay(dt: float, tau: float) -> np.float32:
    return F32(math.exp(float(F32(-F32(dt) / F32(tau)))))
