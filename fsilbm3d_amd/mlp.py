"""Host mirror of the learned warm start's network over the C-ABI
(fsil_mlp_*): a per-marker perceptron 12 -> 64 -> 64 -> 64 -> 3 trained on
the device (FP32 forward/backward/Adam) and evaluated on the matrix cores
(BF16 operands, FP32 accumulation)."""
from __future__ import annotations

from typing import Optional, Tuple

import numpy as np

from . import _lib

NPARAM = 64 * 12 + 64 + 64 * 64 + 64 + 64 * 64 + 64 + 3 * 64 + 3


class WarmStartNet:
    def __init__(self, seed: int = 1, device: int = 0):
        self.lib = _lib.init(device)
        h = np.zeros(1, np.int32)
        _lib.check(self.lib.fsil_mlp_create(int(seed), _lib.iptr(h)))
        self.handle = int(h[0])

    def close(self):
        if getattr(self, "handle", 0):
            _lib.check(self.lib.fsil_mlp_destroy(self.handle))
            self.handle = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def params(self) -> np.ndarray:
        th = np.zeros(NPARAM, np.float32)
        _lib.check(self.lib.fsil_mlp_get_params(self.handle, _lib.fptr(th)))
        return th

    @params.setter
    def params(self, theta: np.ndarray):
        th = np.ascontiguousarray(theta, dtype=np.float32)
        assert th.size == NPARAM
        _lib.check(self.lib.fsil_mlp_set_params(self.handle, _lib.fptr(th)))

    def forward(self, X: np.ndarray, bf16: bool = True) -> np.ndarray:
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.zeros((X.shape[0], 3), np.float32)
        _lib.check(self.lib.fsil_mlp_forward(self.handle, _lib.fptr(X), int(X.shape[0]), _lib.fptr(Y), int(bool(bf16))))
        return Y

    def loss_grad(self, X: np.ndarray, Y: np.ndarray) -> Tuple[float, np.ndarray]:
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.ascontiguousarray(Y, dtype=np.float32)
        loss = np.zeros(1)
        g = np.zeros(NPARAM, np.float32)
        _lib.check(self.lib.fsil_mlp_loss_grad(self.handle, _lib.fptr(X), _lib.fptr(Y), int(X.shape[0]),
                                               _lib.dptr(loss), _lib.fptr(g)))
        return float(loss[0]), g

    def train(self, X: np.ndarray, Y: np.ndarray, epochs: int = 20, batch: int = 1024, lr: float = 1e-3,
              seed: int = 1) -> Tuple[float, float]:
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.ascontiguousarray(Y, dtype=np.float32)
        lb = np.zeros(1)
        la = np.zeros(1)
        _lib.check(self.lib.fsil_mlp_train(self.handle, _lib.fptr(X), _lib.fptr(Y), int(X.shape[0]), int(epochs),
                                           int(batch), float(lr), int(seed), _lib.dptr(lb), _lib.dptr(la)))
        return float(lb[0]), float(la[0])
