"""The warm-start network's CPU oracle against finite differences and the
closed forms of its pieces (no GPU)."""
import numpy as np

from oracle import mlp_oracle as MO


def test_layout_and_size():
    assert MO.NPARAM == 64 * 12 + 64 + 64 * 64 + 64 + 64 * 64 + 64 + 3 * 64 + 3
    th = np.arange(MO.NPARAM, dtype=np.float64)
    assert np.array_equal(MO.pack(MO.unpack(th)), th)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    th = MO.init(11) + 0.01 * rng.standard_normal(MO.NPARAM)
    X = rng.standard_normal((7, 12))
    Y = rng.standard_normal((7, 3))
    _, g = MO.loss_grad(th, X, Y)
    idx = rng.choice(MO.NPARAM, 40, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros(MO.NPARAM)
        e[i] = h
        fd = (MO.loss_grad(th + e, X, Y)[0] - MO.loss_grad(th - e, X, Y)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(fd), (i, fd, g[i])


def test_linear_limit():
    """With tiny weights tanh is the identity to first order: the network is
    the product of its matrices (plus biases)."""
    rng = np.random.default_rng(5)
    th = 1e-4 * rng.standard_normal(MO.NPARAM)
    L = MO.unpack(th)
    X = rng.standard_normal((4, 12))
    W = L[3][0] @ L[2][0] @ L[1][0] @ L[0][0]
    out, _ = MO.forward(th, X)
    lin = X @ W.T + (((L[0][1] @ L[1][0].T + L[1][1]) @ L[2][0].T + L[2][1]) @ L[3][0].T + L[3][1])
    assert np.abs(out - lin).max() < 1e-14


def test_adam_first_step_is_sign_times_lr():
    """Adam's bias correction makes the first step lr * g / (|g| + eps)."""
    g = np.array([0.3, -2.0, 1e-3])
    th, m, v = MO.adam_step(np.zeros(3), g, np.zeros(3), np.zeros(3), 1, 0.01)
    assert np.allclose(th, -0.01 * g / (np.abs(g) + 1e-8), rtol=1e-12)


def test_training_reduces_loss():
    rng = np.random.default_rng(7)
    X = rng.standard_normal((256, 12))
    Y = np.tanh(X[:, :3] * 0.5) * 0.8
    th = MO.init(1)
    m = np.zeros_like(th)
    v = np.zeros_like(th)
    l0 = MO.loss_grad(th, X, Y)[0]
    for t in range(1, 201):
        l, g = MO.loss_grad(th, X, Y)
        th, m, v = MO.adam_step(th, g, m, v, t, 3e-3)
    assert MO.loss_grad(th, X, Y)[0] < 0.1 * l0


def test_init_is_glorot_uniform():
    """Each weight matrix lies within the Glorot-uniform bound sqrt(6/(fan_in+fan_out))
    and fills it (max close to the bound), and every bias is zero."""
    import numpy as np
    from oracle.mlp_oracle import SIZES, init
    th = init(7)
    k = 0
    for o, i in SIZES:
        W = th[k:k + o * i]
        lim = np.sqrt(6.0 / (o + i))
        assert np.abs(W).max() <= lim and np.abs(W).max() > 0.9 * lim
        assert abs(W.mean()) < 0.2 * lim
        assert np.all(th[k + o * i:k + o * i + o] == 0.0)
        k += o * i + o
    assert k == th.size
