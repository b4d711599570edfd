"""Batched symmetric tridiagonal solver on the GPU (DistributedBatchLDLt,
src/processes/pressure.jl:38-201), single device."""

import ctypes

import numpy as np

from . import _lib
from ._lib import BldError, dptr


class BatchLDLt:
    """nk systems of size nz.  ``dvs`` has shape (nk, nz), ``evs`` (nk, nz-1)
    (the reference's γ/β orientation, pressure.jl:67-68); ``precision`` is
    "fp64" or "fp32" (float32 elimination + one float64-residual refinement)."""

    def __init__(self, dvs, evs, precision="fp64"):
        dvs = np.asarray(dvs, dtype=np.float64)
        evs = np.asarray(evs, dtype=np.float64)
        self.nk, self.nz = dvs.shape
        if evs.shape != (self.nk, self.nz - 1):
            raise ValueError("evs must have shape (nk, nz-1)")
        self._lib = _lib.load()
        d = np.ascontiguousarray(dvs.T)  # [nz][nk], mode index fastest
        e = np.ascontiguousarray(evs.T)
        h = ctypes.c_void_p()
        prec = {"fp64": _lib.BLD_FP64, "fp32": _lib.BLD_FP32_MIXED}[precision]
        rc = self._lib.bld_ldlt_create(self.nk, self.nz, dptr(d), dptr(e), prec, ctypes.byref(h))
        if rc != 0:
            raise BldError(rc, self._lib.bld_ldlt_last_error(None).decode())
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._lib.bld_ldlt_destroy(h)
            self._h = None

    def _check(self, rc):
        if rc != 0:
            raise BldError(rc, self._lib.bld_ldlt_last_error(self._h).decode())

    def solve(self, b):
        """x = A \\ b for b of shape (nk, nz), complex."""
        # a copy in [nz][nk] order (the solve works in place; b.T can already be
        # contiguous, for one level or one system, and must not be overwritten)
        buf = np.array(np.asarray(b, dtype=np.complex128).T, order="C", copy=True)
        self._check(self._lib.bld_ldlt_solve(self._h, dptr(buf)))
        return buf.T.copy()

    def bench(self, b, reps=10):
        """(ms per device-resident solve, solution) for b of shape (nk, nz)."""
        buf = np.ascontiguousarray(np.asarray(b, dtype=np.complex128).T)
        x = np.empty_like(buf)
        ms = ctypes.c_double()
        self._check(self._lib.bld_ldlt_bench(self._h, dptr(buf), reps, ctypes.byref(ms), dptr(x)))
        return ms.value, x.T.copy()
