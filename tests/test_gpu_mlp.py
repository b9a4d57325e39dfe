"""The surrogate MLP kernels against the FP64 restatement (oracle/mlp_oracle.py).

Tolerances: the oracle's "bf16" mode rounds at the kernels' rounding points and sums
exactly, so the kernels differ from it only by FP32 summation and by the few BF16 values
that the two round differently -- relative L2 1e-3 on outputs, 5e-3 per gradient tensor.
Against the all-FP64 "exact" mode the BF16 quantisation itself shows: 3e-2 and 5e-2.
"""
import numpy as np
import pytest

import mlp_oracle as mo
from spike_amd import mlp

pytestmark = pytest.mark.gpu
WOUT = (1.0, 0.5, 2.0)


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


def rows(seed, n):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((n, 5)).astype(np.float32)
    t = (np.stack([np.sin(x[:, 0]) + 0.5 * x[:, 1] * x[:, 2], np.tanh(x[:, 3] - x[:, 4]), x[:, 0] * x[:, 4]], 1)
         + 0.05 * rng.standard_normal((n, 3))).astype(np.float32)
    return x, t


@pytest.mark.parametrize("n", [300, 4096])
def test_forward_matches_oracle(n):
    print(f"[mlp] forward 3 members n={n}", flush=True)
    params = mlp.init_params(21, 3)
    m = mlp.Mlp(3, 128)
    m.set_params(params)
    x, _ = rows(22, n)
    y = m.predict(x)
    check = range(n) if n <= 512 else np.random.default_rng(1).choice(n, 300, replace=False)
    check = np.sort(np.asarray(list(check)))
    for e in range(3):
        yb, _ = mo.forward(params[e], x[check], mode="bf16")
        ye, _ = mo.forward(params[e], x[check], mode="exact")
        assert rel(y[e][check], yb) <= 1e-3, e
        assert np.max(np.abs(y[e][check] - yb)) <= 1e-2 * max(1.0, np.max(np.abs(yb)))
        assert rel(y[e][check], ye) <= 3e-2, e
    assert m.timing().forward_ms > 0
    m.close()


@pytest.mark.parametrize("batch", [256, 384])
def test_gradients_match_oracle(batch):
    print(f"[mlp] gradients 2 members batch={batch}", flush=True)
    params = mlp.init_params(31, 2)
    m = mlp.Mlp(2, 512)
    m.set_params(params)
    m.set_loss_weights(WOUT)
    x, t = rows(32, 600)
    m.set_data(x, t)
    loss, g = m.gradients(batch)
    for e in range(2):
        lb, gb, _ = mo.loss_and_grads(params[e], x[:batch], t[:batch], WOUT, mode="bf16")
        le, ge, _ = mo.loss_and_grads(params[e], x[:batch], t[:batch], WOUT, mode="exact")
        assert abs(loss[e] - lb) <= 1e-4 * lb
        for name, (o, s) in mo.OFFSETS.items():
            k = int(np.prod(s))
            assert rel(g[e, o:o + k], gb[o:o + k]) <= 5e-3, (e, name, rel(g[e, o:o + k], gb[o:o + k]))
            assert rel(g[e, o:o + k], ge[o:o + k]) <= 5e-2, (e, name)
    m.close()


def test_adam_and_weight_copies_match_oracle():
    """Three Adam steps from caller gradients; then the BF16 copies the kernels read (W, W^T,
    padded first and last layers) must be those of the updated masters: forward and
    gradients of the updated network are checked against the oracle on the updated
    parameters."""
    print("[mlp] adam", flush=True)
    params = mlp.init_params(41, 2)
    m = mlp.Mlp(2, 256)
    m.set_params(params)
    rng = np.random.default_rng(42)
    opt = [mo.Adam(params[e]) for e in range(2)]
    for step in range(3):
        g = (rng.standard_normal((2, mlp.PARAMS)) * 1e-2).astype(np.float32)
        m.apply_gradients(g, 1e-3)
        for e in range(2):
            opt[e].step(g[e].astype(np.float64), 1e-3)
    p = m.params()
    for e in range(2):
        assert np.max(np.abs(p[e] - opt[e].p)) <= 2e-6, e
    x, t = rows(43, 256)
    m.set_data(x, t)
    y = m.predict(x)
    loss, g = m.gradients(256)
    for e in range(2):
        yb, _ = mo.forward(p[e], x, mode="bf16")
        assert rel(y[e], yb) <= 1e-3
        lb, gb, _ = mo.loss_and_grads(p[e], x, t, mode="bf16")
        for name, (o, s) in mo.OFFSETS.items():
            k = int(np.prod(s))
            assert rel(g[e, o:o + k], gb[o:o + k]) <= 5e-3, (e, name)
    m.close()


def test_full_batch_training_tracks_oracle():
    """20 full-batch Adam steps: the GPU's loss curve follows the oracle's (bf16-mode
    gradients, FP64 Adam) within 2 % at every step, and the loss falls."""
    print("[mlp] full-batch training vs oracle", flush=True)
    params = mlp.init_params(51, 1)
    x, t = rows(52, 256)
    m = mlp.Mlp(1, 256)
    m.set_params(params)
    m.set_data(x, t)
    steps, lr = 20, 1e-3
    lg = m.train(steps, 256, lr, lr, steps, mode=mlp.BATCH_FIRST_ROWS)[:, 0]
    opt = mo.Adam(params[0])
    lo = []
    for s in range(steps):
        l, g, _ = mo.loss_and_grads(opt.p, x, t, mode="bf16")
        lo.append(l)
        opt.step(g, mo.lr_at(s + 1, lr, lr, steps))
    lo = np.array(lo)
    print("gpu", lg[[0, 5, 10, 19]], "oracle", lo[[0, 5, 10, 19]], flush=True)
    assert np.all(np.abs(lg - lo) <= 2e-2 * lo)
    assert lg[-1] < 0.5 * lg[0]
    m.close()


def test_sampled_training_converges_and_graph_replay_is_identical():
    print("[mlp] sampled training, graph vs launches", flush=True)
    x, t = rows(61, 8192)
    params = mlp.init_params(62, 4)
    out = []
    for use_graph in (False, True):
        m = mlp.Mlp(4, 512)
        m.set_params(params)
        m.set_data(x, t)
        loss = m.train(150, 512, 2e-3, 1e-4, 300, mode=mlp.BATCH_SAMPLED, seed=5, record_loss=not use_graph)
        if loss is not None:
            first = loss[:5].mean(0)
        m.train(150, 512, 2e-3, 1e-4, 300, mode=mlp.BATCH_SAMPLED, seed=5, use_graph=use_graph,
                record_loss=False)
        lfin, _ = m.gradients(512)
        out.append((m.params(), lfin))
        m.close()
    (pa, la), (pb, lb) = out
    assert np.array_equal(pa, pb), "graph replay differs from launches"
    print("first", first, "final", la, flush=True)
    assert np.all(la < 0.2 * first)


def test_graph_follows_new_training_data():
    """A captured training step is dropped when the training set (or the loss weights)
    changes, so graph replays after set_data train on the new data exactly as launches do."""
    print("[mlp] graph after new data", flush=True)
    x1, t1 = rows(71, 1024)
    x2, t2 = rows(72, 6000)                       # larger: the data buffers are reallocated
    params = mlp.init_params(73, 2)
    out = []
    for use_graph in (False, True):
        m = mlp.Mlp(2, 256)
        m.set_params(params)
        m.set_data(x1, t1)
        m.train(30, 256, 1e-3, 1e-4, 90, mode=mlp.BATCH_SAMPLED, seed=9, use_graph=use_graph, record_loss=False)
        m.set_data(x2, t2)
        m.train(30, 256, 1e-3, 1e-4, 90, mode=mlp.BATCH_SAMPLED, seed=9, use_graph=use_graph, record_loss=False)
        m.set_loss_weights((1.0, 2.0, 0.5))
        m.train(30, 256, 1e-3, 1e-4, 90, mode=mlp.BATCH_SAMPLED, seed=9, use_graph=use_graph, record_loss=False)
        out.append(m.params())
        m.close()
    assert np.array_equal(out[0], out[1])
