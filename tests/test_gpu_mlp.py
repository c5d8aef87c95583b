"""The warm-start network's device kernels against the CPU oracle
(oracle/mlp_oracle.py): initialisation, FP32 forward, gradients, one Adam
step, BF16 matrix-core inference, and training."""
import numpy as np
import pytest

from oracle import mlp_oracle as MO

pytestmark = pytest.mark.gpu


def _net(seed=3):
    from fsilbm3d_amd.mlp import WarmStartNet
    return WarmStartNet(seed)


def _data(n, seed=0):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, 12)).astype(np.float32)
    Y = (0.5 * np.tanh(X[:, :3]) + 0.1 * X[:, 3:6]).astype(np.float32)
    return X, Y


def test_initialisation_matches_oracle():
    print("[gpu] mlp init", flush=True)
    net = _net(17)
    assert np.array_equal(net.params, MO.init(17).astype(np.float32))


def test_fp32_forward_and_gradient_match_oracle():
    """FP32 device forward within 1e-5 of the float64 oracle; gradients within
    1e-4 relative of the largest gradient (FP32 sums)."""
    print("[gpu] mlp fp32 forward/grad", flush=True)
    net = _net(5)
    th = net.params.astype(np.float64)
    X, Y = _data(300, 1)
    out = net.forward(X, bf16=False)
    ref, _ = MO.forward(th, X.astype(np.float64))
    assert np.abs(out - ref).max() < 1e-5 * max(1.0, np.abs(ref).max())
    loss, g = net.loss_grad(X, Y)
    l_ref, g_ref = MO.loss_grad(th, X.astype(np.float64), Y.astype(np.float64))
    assert abs(loss - l_ref) < 1e-5 * l_ref
    assert np.abs(g - g_ref).max() < 1e-4 * np.abs(g_ref).max()


def test_adam_step_matches_oracle():
    """One full-batch epoch = one Adam step from zero moments."""
    print("[gpu] mlp adam step", flush=True)
    net = _net(9)
    th0 = net.params.astype(np.float64)
    X, Y = _data(128, 2)
    net.train(X, Y, epochs=1, batch=128, lr=1e-3, seed=4)
    _, g = MO.loss_grad(th0, X.astype(np.float64), Y.astype(np.float64))
    th1, _, _ = MO.adam_step(th0, g, np.zeros_like(g), np.zeros_like(g), 1, 1e-3)
    d = net.params.astype(np.float64) - th1
    # the first Adam step moves each parameter by ~lr * sign(g); FP32 sums in
    # another order may flip the sign only where g is ~0
    big = np.abs(g) > 1e-3 * np.abs(g).max()
    assert np.abs(d[big]).max() < 1e-6


def test_bf16_matrix_core_inference_tracks_fp32():
    """BF16 operands with FP32 accumulation: the outputs stay within 2e-2 of
    the float64 oracle (relative to the output scale); a wrong MFMA lane
    layout would give O(1) errors."""
    print("[gpu] mlp bf16 inference", flush=True)
    net = _net(7)
    X, _ = _data(1000, 3)
    out = net.forward(X, bf16=True)
    ref, _ = MO.forward(net.params.astype(np.float64), X.astype(np.float64))
    err = np.abs(out - ref).max() / np.abs(ref).max()
    print(f"[gpu] bf16 rel err {err:.3e}", flush=True)
    assert err < 2e-2


def test_training_reduces_loss():
    print("[gpu] mlp training", flush=True)
    net = _net(11)
    X, Y = _data(4096, 4)
    before, after = net.train(X, Y, epochs=30, batch=256, lr=3e-3, seed=2)
    assert after < 0.05 * before, (before, after)
