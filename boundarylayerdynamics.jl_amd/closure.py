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
