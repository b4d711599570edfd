"""Pins the CPU oracle of the learned SGS closure network."""

import numpy as np

from oracle import mlp as M


def test_bf16_rounding():
    x = np.array([1.0, 1.00390625, 1.005859375, 3.14159265, -2.5e-3, 65504.0])
    r = M.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0  # 1 + 2^-8 ties to even
    assert r[2] == 1.0078125  # above the tie rounds up
    assert abs(r[3] - 3.140625) < 1e-12
    assert np.all(np.abs(r - x) <= np.abs(x) * 2 ** -8)


def test_gelu_grad_matches_finite_differences():
    z = np.linspace(-6, 6, 101)
    h = 1e-6
    fd = (M.gelu(z + h) - M.gelu(z - h)) / (2 * h)
    assert np.allclose(M.gelu_grad(z), fd, atol=1e-8)


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = M.init_params(1)
    for k in p:  # make biases non-zero
        if k.startswith("b"):
            p[k] = 0.1 * rng.standard_normal(p[k].shape)
    p["b3"][M.NOUT:] = 0
    x = np.zeros((20, M.PAD))
    x[:, :M.NIN] = rng.standard_normal((20, M.NIN))
    t = np.zeros((20, M.PAD))
    t[:, :M.NOUT] = rng.standard_normal((20, M.NOUT))
    g = M.backward(p, x, t, round_bf16=False)
    h = 1e-6
    for k in p:
        flat = p[k].ravel()
        idx = rng.choice(flat.size, size=min(12, flat.size), replace=False)
        for i in idx:
            if (k == "W1" and i // M.NH >= M.NIN) or (k == "W3" and i % M.PAD >= M.NOUT) or \
                    (k == "b3" and i >= M.NOUT):
                assert g[k].ravel()[i] == 0
                continue
            old = flat[i]
            flat[i] = old + h
            lp = M.loss(p, x, t, round_bf16=False)
            flat[i] = old - h
            lm = M.loss(p, x, t, round_bf16=False)
            flat[i] = old
            fd = (lp - lm) / (2 * h)
            assert abs(g[k].ravel()[i] - fd) <= 1e-6 * max(1.0, abs(fd)), (k, i, g[k].ravel()[i], fd)


def test_adam_decreases_loss_on_a_fit():
    rng = np.random.default_rng(5)
    x = np.zeros((256, M.PAD))
    x[:, :M.NIN] = rng.standard_normal((256, M.NIN))
    t = np.zeros((256, M.PAD))
    t[:, :M.NOUT] = np.tanh(x[:, :M.NOUT]) * 0.5
    p = M.init_params(2)
    st = {}
    l0 = M.loss(p, x, t)
    for _ in range(60):
        M.adam(p, M.backward(p, x, t), st, lr=3e-3)
    assert M.loss(p, x, t) < 0.5 * l0
