"""CPU oracle of the learned multi-direct-forcing warm start (test
infrastructure only): a 12 -> 64 -> 64 -> 64 -> 3 perceptron with tanh hidden
layers, mean-squared-error loss, backpropagation and Adam, in float64 numpy.

The reference (gaoak/FSILBM3D) has no such network; the method is a learned
warm start added to calculate_interaction_force (Solidbody.f90:869-918):
the network guesses the converged marker force from the interpolated fluid
velocity, the marker velocity, the marker's previous force and its normal.
This file pins the device kernels (forward, gradients, Adam step) against
the textbook formulas; its own gradients are checked by finite differences.

Parameter layout (flat, float32 on the device):
  W1[64][12] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[3][64] b4[3]
"""
from __future__ import annotations

import numpy as np

SIZES = [(64, 12), (64, 64), (64, 64), (3, 64)]
NPARAM = sum(o * i + o for o, i in SIZES)


def unpack(theta):
    theta = np.asarray(theta, dtype=np.float64)
    out, k = [], 0
    for o, i in SIZES:
        W = theta[k:k + o * i].reshape(o, i)
        k += o * i
        b = theta[k:k + o]
        k += o
        out.append((W, b))
    return out


def pack(layers):
    return np.concatenate([np.concatenate([W.reshape(-1), b]) for W, b in layers])


def splitmix64_uniform(seed: int, n: int) -> np.ndarray:
    """n doubles in [0,1): splitmix64 outputs (Steele, Lea & Flood 2014) >> 11,
    times 2^-53; output k uses the state seed + k * 0x9E3779B97F4A7C15."""
    with np.errstate(over="ignore"):
        idx = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed) + idx * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def init(seed: int):
    """Glorot-uniform weights, zero biases, from the splitmix64 stream (entry
    k of the flat layout uses output k+1; biases consume outputs unused)."""
    u = splitmix64_uniform(seed, NPARAM)
    theta = np.zeros(NPARAM)
    k = 0
    for o, i in SIZES:
        lim = np.sqrt(6.0 / (o + i))
        theta[k:k + o * i] = lim * (2.0 * u[k:k + o * i] - 1.0)
        k += o * i + o
    return theta


def forward(theta, X):
    """Returns the output and the hidden activations (for backward)."""
    L = unpack(theta)
    acts = [np.asarray(X, dtype=np.float64)]
    h = acts[0]
    for j, (W, b) in enumerate(L):
        z = h @ W.T + b
        h = np.tanh(z) if j < 3 else z
        acts.append(h)
    return h, acts


def loss_grad(theta, X, Y):
    """MSE over all outputs: mean((f(X) - Y)^2); returns (loss, dloss/dtheta)."""
    L = unpack(theta)
    out, acts = forward(theta, X)
    n = out.size
    d = 2.0 * (out - Y) / n
    grads = [None] * 4
    for j in range(3, -1, -1):
        W, b = L[j]
        hin = acts[j]
        grads[j] = (d.T @ hin, d.sum(axis=0))
        if j > 0:
            dh = d @ W
            d = dh * (1.0 - acts[j] ** 2)
    return float(((out - Y) ** 2).mean()), pack(grads)


def adam_step(theta, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015, algorithm 1); returns new theta, m, v."""
    m = b1 * m + (1.0 - b1) * g
    v = b2 * v + (1.0 - b2) * g * g
    mh = m / (1.0 - b1 ** t)
    vh = v / (1.0 - b2 ** t)
    return theta - lr * mh / (np.sqrt(vh) + eps), m, v
