"""The surrogate MLP on the GPU (include/spike_mlp.h through ctypes).

An ensemble of `models` networks 5 -> 256 x 4 (SiLU) -> 3, trained side by side by the
library's own forward, backward, weight-gradient and Adam kernels (BF16 MFMA, FP32
accumulation, FP32 master weights).  Fails loudly when libspike_hip.so is missing.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import abi

INPUTS, OUTPUTS, HIDDEN, ROW_TILE = 5, 3, 256, 128
BATCH_FIRST_ROWS, BATCH_SAMPLED = 0, 1
SHAPES = [("W1", (HIDDEN, INPUTS)), ("b1", (HIDDEN,)), ("W2", (HIDDEN, HIDDEN)), ("b2", (HIDDEN,)),
          ("W3", (HIDDEN, HIDDEN)), ("b3", (HIDDEN,)), ("W4", (HIDDEN, HIDDEN)), ("b4", (HIDDEN,)),
          ("W5", (OUTPUTS, HIDDEN)), ("b5", (OUTPUTS,))]
PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


class MlpTiming(C.Structure):
    """spk_mlp_timing of include/spike_mlp.h."""
    _fields_ = [("total_ms", C.c_float), ("forward_ms", C.c_float), ("backward_ms", C.c_float),
                ("wgrad_ms", C.c_float), ("update_ms", C.c_float), ("steps", C.c_int64), ("rows", C.c_int64)]


def init_params(seed: int, models: int = 1) -> np.ndarray:
    """He initialisation: W ~ N(0, 2/fan_in) (last layer N(0, 1/fan_in)), biases 0."""
    rng = np.random.default_rng(seed)
    out = np.zeros((models, PARAMS), np.float32)
    for e in range(models):
        parts = []
        for n, s in SHAPES:
            if n.startswith("W"):
                gain = 1.0 if n == "W5" else 2.0
                parts.append((rng.standard_normal(s) * np.sqrt(gain / s[1])).reshape(-1))
            else:
                parts.append(np.zeros(int(np.prod(s))))
        out[e] = np.concatenate(parts).astype(np.float32)
    return out


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class Mlp:
    def __init__(self, models: int, max_batch: int):
        self.lib = abi.load()
        assert self.lib.spk_mlp_param_count() == PARAMS
        h = C.c_void_p()
        abi.check(self.lib.spk_mlp_create(int(models), int(max_batch), C.byref(h)))
        self.h, self.models, self.max_batch = h, int(models), int(max_batch)
        abi._track(self, 1)

    def close(self):
        if getattr(self, "h", None):
            self.lib.spk_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, params: np.ndarray):
        p = _f32(params).reshape(self.models, PARAMS)
        abi.check(self.lib.spk_mlp_set_params(self.h, abi._ptr(p, C.c_float)))

    def params(self) -> np.ndarray:
        out = np.empty((self.models, PARAMS), np.float32)
        abi.check(self.lib.spk_mlp_get_params(self.h, abi._ptr(out, C.c_float)))
        return out

    def reset_optimizer(self):
        abi.check(self.lib.spk_mlp_reset_optimizer(self.h))

    def set_data(self, features: np.ndarray, targets: np.ndarray):
        x, t = _f32(features), _f32(targets)
        assert x.ndim == 2 and x.shape[1] == INPUTS and t.shape == (x.shape[0], OUTPUTS)
        abi.check(self.lib.spk_mlp_set_data(self.h, abi._ptr(x, C.c_float), abi._ptr(t, C.c_float), x.shape[0]))
        self.n = x.shape[0]

    def set_loss_weights(self, w):
        w = _f32(w)
        abi.check(self.lib.spk_mlp_set_loss_weights(self.h, abi._ptr(w, C.c_float)))

    def gradients(self, rows: int):
        g = np.empty((self.models, PARAMS), np.float32)
        loss = np.empty(self.models, np.float32)
        abi.check(self.lib.spk_mlp_gradients(self.h, int(rows), abi._ptr(g, C.c_float), abi._ptr(loss, C.c_float)))
        return loss, g

    def apply_gradients(self, grads: np.ndarray, lr: float):
        g = _f32(grads).reshape(self.models, PARAMS)
        abi.check(self.lib.spk_mlp_apply_gradients(self.h, abi._ptr(g, C.c_float), float(lr)))

    def train(self, steps: int, batch: int, lr0: float, lr_min: float, total_steps: int,
              mode: int = BATCH_SAMPLED, seed: int = 1, use_graph: bool = False,
              record_loss: bool = True) -> Optional[np.ndarray]:
        loss = np.empty((steps, self.models), np.float32) if record_loss else None
        abi.check(self.lib.spk_mlp_train(self.h, int(steps), int(batch), int(mode), int(seed), float(lr0),
                                         float(lr_min), int(total_steps), int(bool(use_graph)),
                                         abi._ptr(loss, C.c_float)))
        return loss

    def predict(self, features: np.ndarray) -> np.ndarray:
        x = _f32(features)
        out = np.empty((self.models, x.shape[0], OUTPUTS), np.float32)
        abi.check(self.lib.spk_mlp_predict(self.h, abi._ptr(x, C.c_float), x.shape[0], abi._ptr(out, C.c_float)))
        return out

    def predict_noalloc(self, features: np.ndarray) -> np.ndarray:
        """predict() through the allocation-free entry point (a fixed 4 MiB staging array
        per GPU; calls overlapping on one GPU race)."""
        x = _f32(features)
        out = np.empty((self.models, x.shape[0], OUTPUTS), np.float32)
        abi.check(self.lib.spk_mlp_predict_noalloc(self.h, abi._ptr(x, C.c_float), x.shape[0],
                                                   abi._ptr(out, C.c_float)))
        return out

    def timing(self) -> MlpTiming:
        t = MlpTiming()
        abi.check(self.lib.spk_mlp_last_timing(self.h, C.byref(t)))
        return t
