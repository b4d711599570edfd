"""Learned SGS closure: the bfloat16-MFMA network behind the C ABI (bld_mlp_*).

The network maps 10 features (9 velocity-gradient components of the filtered
field and the normalised wall distance) to the 6 components of the
deviatoric SGS stress; training and inference run on the GPU.
"""

import ctypes

import numpy as np

from . import _lib
from ._lib import BldError, dptr, fptr

NIN, NH, NOUT, PAD = 10, 64, 6, 16
NPARAM = PAD * NH + NH * NH + NH * PAD + NH + NH + PAD
SLICES = {"W1": (0, (PAD, NH)), "W2": (PAD * NH, (NH, NH)), "W3": (PAD * NH + NH * NH, (NH, PAD)),
          "b1": (PAD * NH + NH * NH + NH * PAD, (NH,)), "b2": (PAD * NH + NH * NH + NH * PAD + NH, (NH,)),
          "b3": (PAD * NH + NH * NH + NH * PAD + 2 * NH, (PAD,))}


def pack(params):
    """dict of arrays (W1 [16,64], W2 [64,64], W3 [64,16], b1, b2, b3) → flat float32."""
    flat = np.zeros(NPARAM, dtype=np.float32)
    for k, (off, shape) in SLICES.items():
        flat[off:off + int(np.prod(shape))] = np.asarray(params[k], dtype=np.float32).ravel()
    return flat


def unpack(flat):
    return {k: np.asarray(flat[off:off + int(np.prod(shape))], dtype=np.float64).reshape(shape)
            for k, (off, shape) in SLICES.items()}


def init_params(seed=0):
    """He-style initialisation with zero padding (input rows ≥ 10, outputs ≥ 6)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    p = {"W1": np.zeros((PAD, NH)), "W2": rng.standard_normal((NH, NH)) * np.sqrt(2.0 / NH),
         "W3": np.zeros((NH, PAD)), "b1": np.zeros(NH), "b2": np.zeros(NH), "b3": np.zeros(PAD)}
    p["W1"][:NIN] = rng.standard_normal((NIN, NH)) * np.sqrt(2.0 / NIN)
    p["W3"][:, :NOUT] = rng.standard_normal((NH, NOUT)) * np.sqrt(1.0 / NH)
    return p


def _pad(a, width=PAD):
    a = np.asarray(a, dtype=np.float32)
    out = np.zeros((a.shape[0], width), dtype=np.float32)
    out[:, :a.shape[1]] = a
    return out


class MLP:
    def __init__(self, params=None, seed=0):
        self._lib = _lib.load()
        flat = pack(init_params(seed) if params is None else params)
        h = ctypes.c_void_p()
        rc = self._lib.bld_mlp_create(fptr(flat), ctypes.byref(h))
        if rc != 0:
            raise BldError(rc, self._lib.bld_mlp_last_error(None).decode())
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._lib.bld_mlp_destroy(h)
            self._h = None

    def _check(self, rc):
        if rc != 0:
            raise BldError(rc, self._lib.bld_mlp_last_error(self._h).decode())

    def params(self):
        flat = np.zeros(NPARAM, dtype=np.float32)
        self._check(self._lib.bld_mlp_get_params(self._h, fptr(flat)))
        return unpack(flat)

    def forward(self, x):
        x = _pad(x)
        y = np.zeros_like(x)
        self._check(self._lib.bld_mlp_forward(self._h, fptr(x), x.shape[0], fptr(y)))
        return y[:, :NOUT]

    def gradient(self, x, t):
        x, t = _pad(x), _pad(t)
        g = np.zeros(NPARAM, dtype=np.float32)
        loss = ctypes.c_double()
        self._check(self._lib.bld_mlp_gradient(self._h, fptr(x), fptr(t), x.shape[0], fptr(g), ctypes.byref(loss)))
        return unpack(g), loss.value

    def set_data(self, x, t):
        self._x, self._t = _pad(x), _pad(t)  # kept alive only for the copy
        self._check(self._lib.bld_mlp_set_data(self._h, fptr(self._x), fptr(self._t), self._x.shape[0]))
        del self._x, self._t

    def train(self, steps, batch=4096, lr=1e-3):
        losses = np.zeros(steps, dtype=np.float64)
        self._check(self._lib.bld_mlp_train(self._h, steps, batch, lr, dptr(losses)))
        return losses

    def time(self, n, reps=10):
        f, g = ctypes.c_double(), ctypes.c_double()
        self._check(self._lib.bld_mlp_time(self._h, n, reps, ctypes.byref(f), ctypes.byref(g)))
        return f.value, g.value


# ---------------------------------------------------------------------------
# training on filtered-DNS samples and the a-priori comparison with Smagorinsky
# ---------------------------------------------------------------------------


def fold_output_scale(params, scale):
    """Parameters whose outputs are `scale` times those of `params` (W3, b3 columns)."""
    p = {k: np.array(v, dtype=np.float64) for k, v in params.items()}
    s = np.zeros(PAD)
    s[:NOUT] = scale
    p["W3"] = p["W3"] * s[None, :]
    p["b3"] = p["b3"] * s
    return p


def train(x, t, steps=2000, batch=8192, lr=2e-3, seed=0, chunks=4):
    """Trains the closure network on the GPU (Adam over consecutive mini-batches
    of the device-resident samples) on targets standardised per component; the
    standardisation is folded into the output layer.  `chunks` splits `steps`
    into runs with a decayed learning rate (lr, lr/3, lr/9, ...).  Returns
    (params dict, losses of the standardised problem)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    order = rng.permutation(x.shape[0])
    x, t = np.asarray(x, dtype=np.float32)[order], np.asarray(t, dtype=np.float32)[order]
    scale = t[:, :NOUT].std(axis=0).astype(np.float64)
    scale[scale == 0] = 1.0
    ts = np.zeros_like(t)
    ts[:, :NOUT] = t[:, :NOUT] / scale[None, :].astype(np.float32)
    net = MLP(seed=seed)
    net.set_data(x, ts)
    losses = []
    per = max(1, steps // chunks)
    for c in range(chunks):
        losses.append(net.train(per, batch=min(batch, x.shape[0]), lr=lr / 3 ** c))
    return fold_output_scale(net.params(), scale), np.concatenate(losses)


def smagorinsky_from_features(x, Cs=0.1, kappa=0.4, damping_exponent=2, wall_damping=True):
    """Static Smagorinsky (static_smagorinsky.jl:125-127, with Mason wall
    damping towards the wall the wall feature measures) written in the
    closure's normalised variables: y = τ^dev/(Δ²|g|²) = −2 (l/Δ)² |Ŝ| Ŝ with
    Ŝ = sym(g/|g|) and l/Δ = (Cs^-N + (κ d/Δ)^-N)^(-1/N), d/Δ = 1/x10 − 1
    (l/Δ = Cs without damping)."""
    g = np.asarray(x[:, :9], dtype=np.float64).reshape(-1, 3, 3)
    S = 0.5 * (g + g.transpose(0, 2, 1))
    mag = np.sqrt(2 * np.einsum("nij,nij->n", S, S))
    xw = np.asarray(x[:, 9], dtype=np.float64)
    with np.errstate(divide="ignore"):
        d = np.where(xw > 0, 1.0 / np.maximum(xw, 1e-300) - 1.0, np.inf)
        N = damping_exponent
        l = (Cs ** (-N) + (kappa * d) ** (-N)) ** (-1.0 / N) if wall_damping else np.full_like(d, Cs)
    tau = -2 * (l ** 2 * mag)[:, None, None] * S
    tr = np.trace(tau, axis1=1, axis2=2) / 3
    tau = tau - tr[:, None, None] * np.eye(3)[None]
    return np.stack([tau[:, 0, 0], tau[:, 1, 1], tau[:, 2, 2], tau[:, 0, 1], tau[:, 0, 2], tau[:, 1, 2]], axis=1)


def apriori(params, x, t, Cs=0.1):
    """Error of the learned stress, of Smagorinsky with constant `Cs` and of
    undamped Smagorinsky with its best constant (least squares over these samples),
    against the filtered-DNS stress, in dimensional units (t[:, 15] = Δ²|ḡ|²):
    relative squared error Σ|τ − τ_true|²/Σ|τ_true|² and the correlation
    coefficient of each component."""
    scale = np.asarray(t[:, 15], dtype=np.float64)[:, None]
    true = np.asarray(t[:, :NOUT], dtype=np.float64) * scale
    w = np.array([1, 1, 1, 2, 2, 2], dtype=np.float64)  # off-diagonal entries count twice in |τ|²

    def metrics(pred):
        err = float(np.sum(w * (pred - true) ** 2) / np.sum(w * true ** 2))
        corr = [float(np.corrcoef(pred[:, m], true[:, m])[0, 1]) if np.std(pred[:, m]) > 0 else 0.0
                for m in range(NOUT)]
        return dict(rel_sq_error=err, corr=corr)

    learned = MLP(params).forward(x).astype(np.float64) * scale
    smag = smagorinsky_from_features(x, Cs) * scale
    m = smagorinsky_from_features(x, 1.0, wall_damping=False) * scale
    c2 = float(np.sum(w * m * true) / np.sum(w * m * m))
    best = m * c2
    return dict(learned=metrics(learned), smagorinsky=metrics(smag), smagorinsky_best=metrics(best),
                best_cs2=c2)
