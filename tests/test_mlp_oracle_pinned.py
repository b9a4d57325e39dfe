"""The FP64 MLP restatement (oracle/mlp_oracle.py) pinned independently of the kernels.

1. Its gradient equals central finite differences of its own loss (exact mode).
2. Its loss and gradient equal torch's CPU autograd in float64 on the same network.
3. Its bf16 rounding equals torch's float32 -> bfloat16 conversion (nearest, ties to even).
4. Its Adam equals torch.optim.Adam step for step; its schedule hits its end points.
5. The bf16 mode stays close to the exact mode and becomes it when nothing needs rounding.
"""
import numpy as np
import pytest
import torch

import mlp_oracle as mo
from spike_amd import mlp

W = (1.0, 0.5, 2.0)


def data(seed, rows=64):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((rows, 5)).astype(np.float32)
    t = rng.standard_normal((rows, 3)).astype(np.float32)
    return x, t


def test_layout_matches_the_product_binding():
    assert mo.P == mlp.PARAMS == 199683
    assert [(n, s) for n, s in mo.SHAPES] == [(n, s) for n, s in mlp.SHAPES]


def test_gradient_equals_finite_differences():
    p = mlp.init_params(3)[0].astype(np.float64)
    x, t = data(4, 32)
    loss, g, _ = mo.loss_and_grads(p, x, t, W, mode="exact")
    rng = np.random.default_rng(5)
    picks = []
    for n, (o, s) in mo.OFFSETS.items():       # parameters of every tensor, biases included
        k = int(np.prod(s))
        picks += list(o + rng.choice(k, size=min(6, k), replace=False))

    def L(q):
        return mo.loss_and_grads(q, x, t, W, mode="exact")[0]

    for i in picks:
        h = 1e-5 * max(1.0, abs(p[i]))
        qp, qm = p.copy(), p.copy()
        qp[i] += h
        qm[i] -= h
        fd = (L(qp) - L(qm)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 2e-5 * abs(g[i]), (i, fd, g[i])


def test_loss_and_gradient_equal_torch_autograd():
    p32 = mlp.init_params(7)[0]
    x, t = data(8, 48)
    loss, g, y = mo.loss_and_grads(p32, x, t, W, mode="exact")
    d = mo.unpack(p32)
    tp = {k: torch.tensor(v, dtype=torch.float64, requires_grad=True) for k, v in d.items()}
    a = torch.tensor(x, dtype=torch.float64)
    for l in range(1, 5):
        a = torch.nn.functional.silu(a @ tp[f"W{l}"].T + tp[f"b{l}"])
    yt = a @ tp["W5"].T + tp["b5"]
    lt = ((torch.tensor(W, dtype=torch.float64) * (yt - torch.tensor(t, dtype=torch.float64)) ** 2).sum()
          / x.shape[0])
    lt.backward()
    gt = mo.pack({k: v.grad.numpy() for k, v in tp.items()})
    assert abs(lt.item() - loss) <= 1e-12 * abs(loss)
    assert np.allclose(yt.detach().numpy(), y, rtol=1e-12, atol=1e-12)
    assert np.max(np.abs(gt - g)) <= 1e-10 * np.max(np.abs(gt))


def test_bf16_rounding_equals_torch():
    rng = np.random.default_rng(9)
    v = np.concatenate([rng.standard_normal(20000).astype(np.float32) * 10.0 ** rng.integers(-20, 20, 20000),
                        np.float32([0.0, -0.0, 1.0, 1.00390625, 1.005859375, 1.0 + 2.0 ** -8,
                                    1.0 + 3 * 2.0 ** -8, 3.0e38, -2.5e-38])]).astype(np.float32)
    ref = torch.tensor(v).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(mo.bf16(v), ref)
    # ties go to even: 1 + 2^-8 lies halfway between 1 and 1 + 2^-7
    assert mo.bf16(np.float32(1.0 + 2.0 ** -8)) == 1.0
    assert mo.bf16(np.float32(1.0 + 3 * 2.0 ** -8)) == 1.0 + 2.0 ** -6


def test_adam_equals_torch_adam_and_schedule_end_points():
    rng = np.random.default_rng(11)
    p0 = rng.standard_normal(500)
    tp = torch.tensor(p0.copy(), requires_grad=True)
    opt = torch.optim.Adam([tp], lr=1e-3, betas=(0.9, 0.999), eps=1e-8)
    ours = mo.Adam(p0)
    for step in range(25):
        g = rng.standard_normal(500) * (1 + step)
        tp.grad = torch.tensor(g)
        opt.step()
        ours.step(g, 1e-3)
    assert np.max(np.abs(ours.p - tp.detach().numpy())) <= 1e-12
    assert mo.lr_at(1, 1e-3, 1e-5, 100) == pytest.approx(1e-3, rel=1e-15)
    assert mo.lr_at(51, 1e-3, 1e-5, 100) == pytest.approx(0.5 * (1e-3 + 1e-5), rel=1e-12)
    assert mo.lr_at(101, 1e-3, 1e-5, 100) == pytest.approx(1e-5, rel=1e-12)
    assert mo.lr_at(500, 1e-3, 1e-5, 100) == pytest.approx(1e-5, rel=1e-12)


def test_bf16_mode_close_to_exact_mode():
    p = mlp.init_params(13)[0]
    x, t = data(14, 128)
    le, ge, ye = mo.loss_and_grads(p, x, t, W, mode="exact")
    lb, gb, yb = mo.loss_and_grads(p, x, t, W, mode="bf16")
    assert np.linalg.norm(yb - ye) <= 2e-2 * np.linalg.norm(ye)
    assert abs(lb - le) <= 2e-2 * le
    for n, (o, s) in mo.OFFSETS.items():
        k = int(np.prod(s))
        assert np.linalg.norm(gb[o:o + k] - ge[o:o + k]) <= 5e-2 * np.linalg.norm(ge[o:o + k]), n
    # with the rounding switched off the two modes are the same computation
    saved = mo.bf16
    try:
        mo.bf16 = lambda v: np.asarray(v, np.float64)
        l2, g2, y2 = mo.loss_and_grads(p, x, t, W, mode="bf16")
    finally:
        mo.bf16 = saved
    assert np.allclose(y2, ye, rtol=1e-13, atol=1e-13) and np.allclose(g2, ge, rtol=1e-11, atol=1e-14)
