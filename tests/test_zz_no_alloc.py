"""The allocation-free LDLt entry point (bld_ldlt_solve_noalloc) against the
handle-based solver (bld_ldlt_create + bld_ldlt_solve) through the C ABI:
the same results on two calls in a row, the second smaller, in float64 and in
float32 with refinement, and a clear error beyond the fixed work space."""

import ctypes

import numpy as np
import pytest

from conftest import load_package

pytestmark = pytest.mark.gpu

B = load_package()
L = B._lib


def pressure_like(nk, nz, seed):
    rng = np.random.default_rng(seed)
    dvs = -(rng.random((nk, nz)) + 2.5)
    evs = rng.random((nk, nz - 1))
    b = rng.standard_normal((nk, nz)) + 1j * rng.standard_normal((nk, nz))
    return dvs, evs, b


def hip():
    """The HIP runtime libbld.so is linked against (already loaded), for the
    test's own device buffers."""
    h = ctypes.CDLL("libamdhip64.so.7", mode=ctypes.RTLD_GLOBAL)
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


H2D, D2H = 1, 2


def to_device(h, a):
    a = np.ascontiguousarray(a)
    p = ctypes.c_void_p()
    assert h.hipMalloc(ctypes.byref(p), a.nbytes) == 0
    assert h.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, H2D) == 0
    return p


def noalloc_solve(dvs, evs, b, precision):
    lib, h = L.load(), hip()
    nk, nz = dvs.shape
    d = to_device(h, dvs.T)  # [nz][nk]
    e = to_device(h, evs.T if nz > 1 else np.zeros((1, nk)))
    xb = np.ascontiguousarray(b.T)  # [nz][nk] complex, interleaved
    x = to_device(h, xb)
    prec = {"fp64": L.BLD_FP64, "fp32": L.BLD_FP32_MIXED}[precision]
    try:
        rc = lib.bld_ldlt_solve_noalloc(nk, nz, d, e, x, prec, None)
        assert rc == 0, lib.bld_ldlt_noalloc_last_error().decode()
        assert h.hipDeviceSynchronize() == 0
        out = np.empty_like(xb)
        assert h.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), x, out.nbytes, D2H) == 0
    finally:
        for p in (d, e, x):
            h.hipFree(p)
    return out.T.copy()


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-14), ("fp32", 1e-7)])
def test_noalloc_matches_handle_solver(precision, tol):
    for nk, nz, seed in ((1500, 300, 1), (37, 17, 2)):  # second call smaller
        dvs, evs, b = pressure_like(nk, nz, seed)
        want = B.BatchLDLt(dvs, evs, precision=precision).solve(b)
        got = noalloc_solve(dvs, evs, b, precision)
        err = np.abs(got - want).max() / np.abs(want).max()
        assert err <= tol, (precision, nk, nz, err)


def test_noalloc_single_level_and_limits():
    dvs, evs, b = pressure_like(10, 1, 3)
    want = B.BatchLDLt(dvs, evs).solve(b)
    assert np.abs(noalloc_solve(dvs, evs, b, "fp64") - want).max() <= 1e-15 * np.abs(want).max()
    lib = L.load()
    one = ctypes.c_void_p(1)  # never dereferenced: the sizes are rejected first
    rc = lib.bld_ldlt_solve_noalloc(65536, 256, one, one, one, L.BLD_FP64, None)
    assert rc == L.BLD_EARG and b"more than the fixed" in lib.bld_ldlt_noalloc_last_error()
    rc = lib.bld_ldlt_solve_noalloc(4, 5000, one, one, one, L.BLD_FP64, None)
    assert rc == L.BLD_EARG and b"4096 levels" in lib.bld_ldlt_noalloc_last_error()
