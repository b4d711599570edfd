"""The bfloat16-MFMA closure network against its CPU oracle (oracle/mlp.py)."""

import numpy as np
import pytest

from oracle import mlp as M
from conftest import load_package

pytestmark = pytest.mark.gpu
B = load_package()
C = B.closure


def data(n, seed):
    rng = np.random.default_rng(seed)
    x = np.zeros((n, M.PAD))
    x[:, :M.NIN] = rng.standard_normal((n, M.NIN))
    t = np.zeros((n, M.PAD))
    t[:, :M.NOUT] = rng.standard_normal((n, M.NOUT))
    return x, t


@pytest.mark.parametrize("n", [16, 37, 1000])
def test_forward_matches_oracle(n):
    p = M.init_params(4)
    for k in ("b1", "b2"):
        p[k] = 0.1 * np.random.default_rng(1).standard_normal(p[k].shape)
    net = C.MLP(p)
    x, _ = data(n, n)
    y = net.forward(x[:, :M.NIN])
    yo, _ = M.forward(p, x)
    # float32 accumulation vs float64: differences from the order of the sums and
    # from bf16 re-rounding of activations that land near a rounding boundary
    err = np.abs(y - yo[:, :M.NOUT]).max() / np.abs(yo[:, :M.NOUT]).max()
    assert err < 2e-2, err
    assert np.median(np.abs(y - yo[:, :M.NOUT])) < 1e-3 * np.abs(yo).max()


def test_gradient_matches_oracle():
    p = M.init_params(5)
    net = C.MLP(p)
    x, t = data(2048, 7)
    g, loss = net.gradient(x[:, :M.NIN], t[:, :M.NOUT])
    go = M.backward(p, x, t)
    lo = M.loss(p, x, t)
    assert abs(loss - lo) < 1e-3 * lo
    for k in go:
        a, b = g[k].ravel(), go[k].ravel()
        cos = a @ b / (np.linalg.norm(a) * np.linalg.norm(b))
        assert cos > 0.999, (k, cos)
        assert np.linalg.norm(a - b) < 2e-2 * np.linalg.norm(b), (k, np.linalg.norm(a - b) / np.linalg.norm(b))


def test_training_reduces_loss():
    rng = np.random.default_rng(9)
    n = 16384
    x = rng.standard_normal((n, M.NIN))
    t = np.stack([np.tanh(x[:, i]) * 0.5 + 0.2 * x[:, (i + 1) % M.NIN] for i in range(M.NOUT)], axis=1)
    net = C.MLP(seed=3)
    net.set_data(x, t)
    losses = net.train(200, batch=4096, lr=3e-3)
    assert np.all(np.isfinite(losses))
    assert losses[-10:].mean() < 0.2 * losses[:3].mean(), (losses[:3], losses[-10:])
