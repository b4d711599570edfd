"""CPU oracle of the learned SGS closure network (test infrastructure only).

The closure has no counterpart in the reference: it is an addition
that sits behind the reference's process API next to StaticSmagorinskyModel
(static_smagorinsky.jl:39-42).  This module restates the network the GPU
kernels evaluate, in float64 NumPy:

    x (10 features, padded to 16) → h1 = GELU(x·W1 + b1) (64)
                                  → h2 = GELU(h1·W2 + b2) (64)
                                  → y  = h2·W3 + b3        (6, padded to 16)

with GELU(z) = z/2 · (1 + tanh(√(2/π)·(z + 0.044715 z³))).  Matrix operands
(x, h1, h2 and the weights) are rounded to bfloat16 exactly as the MFMA
kernels round them; sums, biases and activations stay in float64 here.  The
analytic backward pass is pinned by central finite differences (test_mlp_*).
"""

import numpy as np

NIN, NH, NOUT, PAD = 10, 64, 6, 16
SQ = np.sqrt(2.0 / np.pi)


def bf16(a):
    """Round float32/float64 values to bfloat16 (round to nearest even), as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(SQ * (z + 0.044715 * z ** 3)))


def gelu_grad(z):
    t = np.tanh(SQ * (z + 0.044715 * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t ** 2) * SQ * (1.0 + 3 * 0.044715 * z ** 2)


def init_params(seed=0):
    """Padded parameter set: W1 [16,64], W2 [64,64], W3 [64,16], b1 [64], b2 [64], b3 [16]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    W1 = np.zeros((PAD, NH))
    W1[:NIN] = rng.standard_normal((NIN, NH)) * np.sqrt(2.0 / NIN)
    W2 = rng.standard_normal((NH, NH)) * np.sqrt(2.0 / NH)
    W3 = np.zeros((NH, PAD))
    W3[:, :NOUT] = rng.standard_normal((NH, NOUT)) * np.sqrt(1.0 / NH)
    return dict(W1=W1, W2=W2, W3=W3, b1=np.zeros(NH), b2=np.zeros(NH), b3=np.zeros(PAD))


def forward(p, x, round_bf16=True):
    """x [N, 16] → (y [N, 16], cache)."""
    r = bf16 if round_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    xb = r(x)
    z1 = xb @ r(p["W1"]) + p["b1"]
    h1 = r(gelu(z1))
    z2 = h1 @ r(p["W2"]) + p["b2"]
    h2 = r(gelu(z2))
    y = h2 @ r(p["W3"]) + p["b3"]
    return y, dict(xb=xb, z1=z1, h1=h1, z2=z2, h2=h2)


def loss(p, x, t, round_bf16=True):
    """Mean squared error over the points and the 6 outputs."""
    y, _ = forward(p, x, round_bf16)
    d = (y - t)[:, :NOUT]
    return np.sum(d * d) / (x.shape[0] * NOUT)


def backward(p, x, t, round_bf16=True):
    """Gradient of `loss` with respect to every parameter (float64)."""
    r = bf16 if round_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    y, c = forward(p, x, round_bf16)
    n = x.shape[0]
    dy = np.zeros_like(y)
    dy[:, :NOUT] = 2.0 * (y - t)[:, :NOUT] / (n * NOUT)
    g = {}
    g["W3"] = c["h2"].T @ dy
    g["b3"] = dy.sum(0)
    dz2 = (dy @ r(p["W3"]).T) * gelu_grad(c["z2"])
    g["W2"] = c["h1"].T @ dz2
    g["b2"] = dz2.sum(0)
    dz1 = (dz2 @ r(p["W2"]).T) * gelu_grad(c["z1"])
    g["W1"] = c["xb"].T @ dz1
    g["b1"] = dz1.sum(0)
    g["W1"][NIN:] = 0  # padded input rows and output columns carry no parameters
    g["W3"][:, NOUT:] = 0
    g["b3"][NOUT:] = 0
    return g


def adam(p, g, state, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015) on float64 master parameters."""
    state["t"] = state.get("t", 0) + 1
    t = state["t"]
    for k in p:
        m = state.setdefault("m_" + k, np.zeros_like(p[k]))
        v = state.setdefault("v_" + k, np.zeros_like(p[k]))
        m[...] = beta1 * m + (1 - beta1) * g[k]
        v[...] = beta2 * v + (1 - beta2) * g[k] ** 2
        mh = m / (1 - beta1 ** t)
        vh = v / (1 - beta2 ** t)
        p[k] = p[k] - lr * mh / (np.sqrt(vh) + eps)
    return p
