"""Network construction: the oracle's restatement of Spike's front-end construction
(oracle/construction.py) against the fixtures the reference front-end wrote
(tests/golden/*, tests/golden/gen_connectivity.cpp).  Every case below builds the same
network, call for call, as the matching function of gen_connectivity.cpp."""
import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
import construction as C  # noqa: E402

GOLD = os.path.join(ROOT, "tests", "golden")
F = np.float32


def load(case):
    d = os.path.join(GOLD, case)
    meta = json.load(open(os.path.join(d, "meta.json")))

    def arr(name, dt):
        return np.fromfile(os.path.join(d, name + ".bin"), dtype=dt)

    return meta, arr


def brunel_small():
    b = C.Builder(42)
    inp = b.add_inputs(100)
    e = b.add_neurons(80)
    i = b.add_neurons(20)
    rule = b.add_rule("wd")
    wv = F(F(0.1) * F(np.power(F(10.0), F(-3.0))))
    delay = F(F(1.5) * F(np.power(F(10.0), F(-3.0))))
    ex = dict(weight_range=(wv, wv), delay_range=(delay, delay))
    ih = dict(weight_range=(F(-5.0) * wv, F(-5.0) * wv), delay_range=(delay, delay))
    b.connect_with_sparsity(inp, e, 100, 80, 0.1, **ex)
    b.connect_with_sparsity(inp, i, 100, 20, 0.1, **ex)
    b.connect_with_sparsity(e, i, 80, 20, 0.1, **ex)
    b.connect_with_sparsity(e, e, 80, 80, 0.1, rules=(rule,), **ex)
    b.connect_with_sparsity(i, e, 20, 80, 0.1, **ih)
    b.connect_with_sparsity(i, i, 20, 20, 0.1, **ih)
    return b


def plain_small():
    b = C.Builder(42, conductance=True)
    inp = b.add_inputs(100)
    n = b.add_neurons(100)
    b.add_group(inp, n, kind="random", p=0.1, weight_range=(0.0, 1.0), scale=1e-9,
                delay_range=(0.0001, 0.0010), vhat=0.0, tau_g=0.001)
    return b


def coba_small():
    b = C.Builder(7, conductance=True)
    inp = b.add_inputs(40)
    e = b.add_neurons(80)
    i = b.add_neurons(20)
    vog = b.add_rule("vogels")
    ev = b.add_rule("evans")
    common = dict(kind="random", p=0.1, scale=1e-9, delay_range=(0.0008, 0.0008))
    b.add_group(inp, e, weight_range=(0.2, 0.8), vhat=0.0, tau_g=0.005, rules=(ev,), **common)
    b.add_group(e, e, weight_range=(0.6, 0.6), vhat=0.0, tau_g=0.005, **common)
    b.add_group(e, i, weight_range=(0.6, 0.6), vhat=0.0, tau_g=0.005, **common)
    b.add_group(i, e, weight_range=(6.7, 6.7), vhat=-0.080, tau_g=0.010, rules=(vog,), **common)
    b.add_group(i, i, weight_range=(6.7, 6.7), vhat=-0.080, tau_g=0.010, **common)
    return b


def mixed_small():
    b = C.Builder(3)
    in0 = b.add_inputs(10)
    in1 = b.add_inputs(6)
    a = b.add_neurons(20)
    bb = b.add_neurons(20)
    allp = dict(kind="all", weight_range=(0.5e-3, 2.0e-3), delay_range=(0.0002, 0.0020))
    b.add_group(in0, a, **allp)
    b.add_group(in1, bb, **allp)
    b.add_group(a, bb, kind="one", weight_range=(1.0e-3, 1.0e-3), scale=2.0, delay_range=(0.0003, 0.0003))
    b.add_group(bb, a, **allp)
    return b


CASES = {"brunel_ref_small": brunel_small, "plain_small": plain_small,
         "coba_small": coba_small, "mixed_small": mixed_small}


def test_glibc_rand_known_answers():
    # glibc srand(1): the first outputs of rand() (documented values of the TYPE_3 generator)
    r = C.GlibcRand(1)
    assert [r.rand() for _ in range(5)] == [1804289383, 846930886, 1681692777, 1714636915, 1957747793]
    # srand(0) is srand(1)
    r0 = C.GlibcRand(0)
    assert r0.rand() == 1804289383


@pytest.mark.parametrize("case", sorted(CASES))
def test_creation_order_matches_reference(case):
    meta, arr = load(case)
    b = CASES[case]()
    assert len(b.pre) == meta["n_synapses"]
    assert b.N == meta["n_neurons"] and b.Nin == meta["n_inputs"]
    np.testing.assert_array_equal(np.array(b.pre, np.int32), arr("orig_pre", np.int32))
    np.testing.assert_array_equal(np.array(b.post, np.int32), arr("orig_post", np.int32))
    np.testing.assert_array_equal(np.array(b.w, np.float32).view(np.uint32), arr("orig_w", np.float32).view(np.uint32))
    np.testing.assert_array_equal(np.array(b.scale, np.float32), arr("orig_scale", np.float32))
    np.testing.assert_array_equal(np.array(b.delay, np.int32), arr("orig_delay", np.int32))
    np.testing.assert_array_equal(np.array(b.label, np.int32), arr("orig_label", np.int32))


@pytest.mark.parametrize("case", sorted(CASES))
def test_sorted_csr_matches_reference(case):
    meta, arr = load(case)
    out = CASES[case]().finalise()
    for k in ["pre", "post", "delay", "label"]:
        np.testing.assert_array_equal(out[k], arr("sorted_" + k, np.int32), err_msg=k)
    for k in ["w", "scale"]:
        np.testing.assert_array_equal(out[k].view(np.uint32), arr("sorted_" + k, np.float32).view(np.uint32), err_msg=k)
    np.testing.assert_array_equal(out["sort_idx"], arr("sort_idx", np.int32))
    np.testing.assert_array_equal(out["reversesort_idx"], arr("reversesort_idx", np.int32))
    for k in ["neuron_eff_start", "neuron_eff_count", "input_eff_start", "input_eff_count"]:
        got = arr(k, np.int32)
        exp = out[k]
        if k.endswith("start"):
            # starts are only written for neurons with efferent synapses
            cnt = out[k.replace("start", "count")]
            np.testing.assert_array_equal(exp[cnt > 0], got[cnt > 0], err_msg=k)
        else:
            np.testing.assert_array_equal(exp, got, err_msg=k)
    assert out["timestep_grouping"] == meta["timestep_grouping"]
    assert out["min_delay"] == meta["min_delay"] and out["max_delay"] == meta["max_delay"]
    assert out["num_syn_labels"] == meta["num_syn_labels"]
    for r, rule in enumerate(meta["rules"]):
        np.testing.assert_array_equal(out["plastic"][r], arr("plastic_%d" % r, np.int32))
        assert rule["n"] == len(out["plastic"][r])


def test_brunel_fixture_invariants():
    """Brute-force counts of the fixed in-degree rule: every LIF neuron receives exactly
    0.1 x (group size) synapses from each source population."""
    meta, arr = load("brunel_ref_small")
    post = arr("orig_post", np.int32)
    pre = arr("orig_pre", np.int32)
    indeg_in = np.bincount(post[pre < 0], minlength=100)
    assert np.all(indeg_in == 10)
    from_e = (pre >= 0) & (pre < 80)
    from_i = pre >= 80
    assert np.all(np.bincount(post[from_e], minlength=100) == 8)
    assert np.all(np.bincount(post[from_i], minlength=100) == 2)
