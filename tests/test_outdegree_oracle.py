"""Pins of the oracle's fixed out-degree generator (or_outdegree_connectivity), the rule the
library draws the large plastic network's synapses with on the device.

Rows are recomputed draw by draw from the Philox4x32-10 block (itself pinned by the
Random123 known-answer vectors in test_oracle_pinned.py); the row layout, degrees and
ranges are counted by brute force; the target, delay and weight distributions are
checked against the uniform laws they must follow.  CPU only."""
import numpy as np
import pytest
from scipy import stats

import oracle_py
from spike_amd import builders
from spike_amd.netdesc import OutDegree

N, NIN, NE, K, KIN, DMAX = 1000, 60, 800, 50, 40, 20
SEED = (7 << 32) | 12345     # both key words non-zero
O = OutDegree(n_exc=NE, out_degree=K, input_out_degree=KIN, delay_max=DMAX, scale_exc=1e-4,
              scale_inh=-5e-4, scale_input=2e-3, weight_inh=1.0, weight_input=0.75)


@pytest.fixture(scope="module")
def arrays():
    return oracle_py.outdegree_connectivity(SEED, N, NIN, O, threads=4)


def test_rows_recomputed_from_philox(arrays):
    pre, post, w, sc, dl = arrays
    key = [SEED & 0xFFFFFFFF, SEED >> 32]
    for u in (0, 1, NE - 1, NE, N - 1, N, N + NIN - 1):
        base = u * K if u <= N else N * K + (u - N) * KIN
        for j in range(K if u < N else KIN):
            r = [int(x) for x in oracle_py.philox([j, 0, u, 5], key)]
            s = base + j
            assert pre[s] == (u if u < N else -(u - N) - 1)
            assert post[s] == (r[0] * N) >> 32
            assert dl[s] == 1 + ((r[1] * DMAX) >> 32)
            if u < NE:
                assert w[s] == np.float32((r[2] >> 8) * 2.0 ** -24) and sc[s] == np.float32(1e-4)
            elif u < N:
                assert w[s] == 1.0 and sc[s] == np.float32(-5e-4)
            else:
                assert w[s] == np.float32(0.75) and sc[s] == np.float32(2e-3)


def test_degrees_layout_and_ranges(arrays):
    pre, post, w, sc, dl = arrays
    assert len(pre) == N * K + NIN * KIN
    # rows are contiguous and in unit order: neurons 0..N-1, then inputs
    unit = np.where(pre >= 0, pre, N + (-pre - 1))
    assert np.all(np.diff(unit) >= 0)
    counts = np.bincount(unit, minlength=N + NIN)
    assert np.all(counts[:N] == K) and np.all(counts[N:] == KIN)
    assert post.min() >= 0 and post.max() < N
    assert dl.min() >= 1 and dl.max() <= DMAX
    we = w[:NE * K]
    assert we.min() >= 0.0 and we.max() < 1.0
    assert np.all(we * np.float32(2 ** 24) == np.floor(we * np.float32(2 ** 24)))   # multiples of 2^-24


def test_target_delay_weight_laws(arrays):
    pre, post, w, sc, dl = arrays
    S = len(post)
    # targets uniform over the neurons, delays uniform over 1..DMAX (chi-square)
    assert stats.chisquare(np.bincount(post, minlength=N)).pvalue > 1e-4
    assert stats.chisquare(np.bincount(dl, minlength=DMAX + 1)[1:]).pvalue > 1e-4
    # excitatory weights U[0, 1): Kolmogorov-Smirnov
    assert stats.kstest(w[:NE * K].astype(np.float64), "uniform").pvalue > 1e-4
    # no dependence between a synapse's target and its delay (different outputs of the block)
    assert abs(np.corrcoef(post, dl)[0, 1]) < 5.0 / np.sqrt(S)


def test_thread_count_does_not_change_the_arrays(arrays):
    one = oracle_py.outdegree_connectivity(SEED, N, NIN, O, threads=1)
    for a, b in zip(arrays, one):
        assert np.array_equal(a, b)


def test_expanded_network_and_its_rule():
    d = builders.large_plastic(n=2000, out_degree=30, input_out_degree=20, seed=5)
    e = oracle_py.expand_outdegree(d, threads=2)
    o = d.outdegree
    assert e.outdegree is None and e.n_synapses == d.n_synapses == 2000 * 30 + 100 * 20
    assert e.timestep_grouping() == 1
    ee = e.rules[0].synapses
    brute = [s for s in range(e.n_synapses) if 0 <= e.pre[s] < o.n_exc and e.post[s] < o.n_exc]
    assert np.array_equal(ee, np.array(brute, np.int64))
    assert e.rules[0].params == o.ee_stdp
    # the inhibitory and input scales are the float32 values the host builder uses
    assert np.float32(o.scale_inh) == np.float32(-5.0 * 0.1e-3)
    assert set(np.unique(e.scale).tolist()) == {float(np.float32(0.1e-3)), float(np.float32(-5.0 * 0.1e-3)),
                                                float(np.float32(2.0e-3))}
