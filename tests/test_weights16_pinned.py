"""16-bit weight and trace storage of the FP32 oracle, pinned independently.

1. The oracle's rounding to bfloat16 equals torch's float32 -> bfloat16 conversion and its
   rounding to binary16 equals numpy's float32 -> float16 conversion (nearest, ties to even,
   subnormals, overflow to infinity), bit for bit; decoding is exact.
2. Kahan-compensated storage keeps small weight updates: with Vogels iSTDP at a learning
   rate whose every per-group change stays below half a bfloat16 step (so rounding each new
   weight to bfloat16 alone would never move a weight), the BF16-stored weights move as far
   as the FP32 ones: several steps, ending within a quarter of a step of them (median) and
   one at most.  (The plastic synapses carry no current, so both runs see the same spikes;
   the traces are binary16 in both 16-bit modes.)
"""
import numpy as np
import torch

import oracle_py
from spike_amd import builders
from spike_amd.netdesc import Rule, RULE_VOGELS, WEIGHTS_BF16, WEIGHTS_FP16


def values():
    rng = np.random.default_rng(3)
    v = [rng.standard_normal(4000) * 10.0 ** rng.integers(-9, 6, 4000),
         rng.uniform(-1, 1, 2000) * 6.2e-5,                      # binary16 subnormal range
         np.float32([0.0, -0.0, 1.0, 65504.0, 65519.0, 65520.0, 1e9, -1e9, 2.0 ** -24, 2.0 ** -25,
                     3 * 2.0 ** -26, 2.0 ** -14, 1.0 + 2.0 ** -11, 1.0 + 3 * 2.0 ** -11,
                     1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8])]
    return np.concatenate(v).astype(np.float32)


def test_rounding_equals_torch_and_numpy():
    lib = oracle_py.load()
    x = values()
    bf = np.array([lib.or_round16_bits(float(v), WEIGHTS_BF16) for v in x], np.uint16)
    hf = np.array([lib.or_round16_bits(float(v), WEIGHTS_FP16) for v in x], np.uint16)
    ref_bf = torch.tensor(x).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    with np.errstate(over="ignore"):
        ref_hf = x.astype(np.float16).view(np.uint16)
        ref_hf_dec = x.astype(np.float16).astype(np.float32)
    assert np.array_equal(bf, ref_bf)
    assert np.array_equal(hf, ref_hf)
    dec_bf = np.array([lib.or_decode16(int(h), WEIGHTS_BF16) for h in bf], np.float32)
    dec_hf = np.array([lib.or_decode16(int(h), WEIGHTS_FP16) for h in hf], np.float32)
    assert np.array_equal(dec_bf, torch.tensor(x).to(torch.bfloat16).float().numpy())
    assert np.array_equal(dec_hf, ref_hf_dec)


def _vogels_net(fmt):
    """Drive through the random network's synapses; a plastic copy of every synapse with
    scale 0 (so the plastic weights never change the spikes, and every storage format sees
    the same spike trains), all plastic weights starting at 2^-10 (exact in bfloat16)."""
    d = builders.random_net(seed=21, n_neurons=64, n_inputs=64, p=0.3, input_rate=400.0, whi=3e-3,
                            dmin=1, dmax=4)
    S = d.n_synapses
    sc = d.scale if d.scale is not None else np.ones(S, np.float32)
    d.pre = np.concatenate([d.pre, d.pre])
    d.post = np.concatenate([d.post, d.post])
    d.delay = np.concatenate([d.delay, d.delay])
    d.weight = np.concatenate([d.weight, np.full(S, np.float32(2.0 ** -10), np.float32)])
    d.scale = np.concatenate([sc, np.zeros(S, np.float32)]).astype(np.float32)
    if d.label is not None:
        d.label = np.concatenate([d.label, d.label])
    syn = np.arange(S, 2 * S, dtype=np.int64)
    d.rules.append(Rule(RULE_VOGELS, syn, builders.vogels_params(lr=1.2e-7, rate=10.0)))
    d.weight_format = fmt
    return d


def test_kahan_storage_keeps_small_updates():
    w0 = 2.0 ** -10
    ulp = 2.0 ** -17                                      # bfloat16 step in [2^-10, 2^-9)
    groups = 1500
    # FP32 storage, group by group: every group's change is far below half a step, so a
    # bfloat16 weight rounded after each update would never leave 1e-3
    d = _vogels_net(0)
    G = d.timestep_grouping()
    o = oracle_py.OracleNetwork(d, precision=32)
    plastic = d.rules[0].synapses
    w_prev = o.weights()[plastic]
    step_max = 0.0
    for s in range(groups):
        o.run(float(d.group_time(0.0, s)), 1)
        w = o.weights()[plastic]
        step_max = max(step_max, float(np.max(np.abs(w - w_prev))))
        w_prev = w
    w32 = w_prev
    o.close()
    assert step_max < 0.5 * ulp, step_max / ulp
    assert np.median(np.abs(w32 - w0)) > 2 * ulp, "the weights must move by several bfloat16 steps"
    # BF16 storage with Kahan compensation follows the FP32 weights
    d = _vogels_net(WEIGHTS_BF16)
    o = oracle_py.OracleNetwork(d, precision=32)
    o.run(0.0, groups)
    w16 = o.weights()[plastic]
    o.close()
    diff = np.abs(w16 - w32)
    assert np.all(w32 > w0)                               # the weights grow: steps of 2^-17
    assert np.median(diff) <= 0.25 * ulp and np.max(diff) <= ulp, (np.median(diff) / ulp, np.max(diff) / ulp)
