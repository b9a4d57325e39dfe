"""spk_mlp_predict_noalloc (the allocation-free variant of spk_mlp_predict) through the C
ABI: its outputs equal spk_mlp_predict's bit for bit on two calls in a row, the second
with fewer rows than the first (the shared staging array keeps the first call's
contents), with another handle's call in between; sizes beyond the fixed 4 MiB staging
array are refused with SPK_ERR_ARG before anything runs."""
import numpy as np
import pytest

import mlp_oracle as mo
from spike_amd import abi, mlp

pytestmark = pytest.mark.gpu

MAX_ROWS_8 = 36157   # round_up(5 n, 64) + 24 n <= 1,048,576 floats


def test_noalloc_predict_equals_predict():
    print("[no-alloc] 8 members: 36157 rows, then 1000 rows", flush=True)
    params = mlp.init_params(41, 8)
    m = mlp.Mlp(8, 128)
    m.set_params(params)
    other = mlp.Mlp(2, 128)
    other.set_params(mlp.init_params(43, 2))
    rng = np.random.default_rng(5)
    for n in (MAX_ROWS_8, 1000):
        x = rng.standard_normal((n, 5)).astype(np.float32)
        got = m.predict_noalloc(x)
        ref = m.predict(x)
        assert got.shape == (8, n, 3)
        assert np.array_equal(got, ref), n
        assert m.timing().forward_ms > 0
        # the kernel's outputs themselves agree with the FP64 restatement (bf16 mode)
        yb, _ = mo.forward(params[3], x[:64], mode="bf16")
        assert float(np.linalg.norm(got[3][:64] - yb) / np.linalg.norm(yb)) <= 1e-3
        # a call on another handle uses the same array in between
        xo = rng.standard_normal((777, 5)).astype(np.float32)
        assert np.array_equal(other.predict_noalloc(xo), other.predict(xo))
    m.close()
    other.close()


def test_noalloc_predict_rejects_sizes_beyond_the_array():
    print("[no-alloc] size limit", flush=True)
    m = mlp.Mlp(8, 128)
    m.set_params(mlp.init_params(47, 8))
    with pytest.raises(abi.SpikeError, match="static limit"):
        m.predict_noalloc(np.zeros((MAX_ROWS_8 + 1, 5), np.float32))
    # the handle stays usable
    x = np.random.default_rng(9).standard_normal((200, 5)).astype(np.float32)
    assert np.array_equal(m.predict_noalloc(x), m.predict(x))
    m.close()
