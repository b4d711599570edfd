"""CPU-side checks of the C ABI library and the host discretisation (no GPU calls)."""

import ctypes
import os
import re

import numpy as np
import pytest

import oracle as O
from conftest import REPO, load_package

B = load_package()


def test_library_exports_every_declared_symbol():
    lib = B._lib.load()
    header = open(os.path.join(REPO, "include", "bld.h")).read()
    declared = set(re.findall(r"\b(bld_\w+)\s*\(", header))
    assert declared, "no declarations found"
    assert declared == set(B._lib.EXPORTS)
    for name in declared:
        assert getattr(lib, name) is not None


def test_config_struct_matches_header_layout():
    # 4 ints, 2 ints, 2 doubles, 2 pointers, 6 bcs (24 B), ... computed by the C compiler rule
    c = B._lib.BldConfig
    assert c.bc.offset % 8 == 0 and c.sgs_lc.offset % 8 == 0
    assert ctypes.sizeof(B._lib.BldBC) == 24


CFGS = [
    ((128, 128, 64), (2 * np.pi, 2 * np.pi, 1.0), ("RoughWall", 1e-4), "FreeSlipBoundary", None),
    ((64, 64, 64), (2 * np.pi, np.pi, 2.0), "FreeSlipBoundary", "FreeSlipBoundary", None),
    ((16, 12, 9), (1.0, 2.0, 1.0), ("RoughWall", 1e-3), ("RoughWall", 1e-3), 0.6),
]


@pytest.mark.parametrize("res,dims,lo,hi,eta", CFGS)
def test_discretization_constants_match_oracle(res, dims, lo, hi, eta):
    def mk(M):
        b = lambda s: getattr(M, s)() if isinstance(s, str) else getattr(M, s[0])(s[1])
        mapping = None if eta is None else M.SinusoidalMapping(eta)
        return M.Domain(dims, b(lo), b(hi), mapping)

    procs = lambda M: M.incompressible_flow(1e-4, sgs_model=M.StaticSmagorinskyModel())
    d = B.model.Discretization(res, mk(B), procs(B))
    ref = O.Model(res, mk(O), procs(O))
    c = d.config
    n3 = res[2]
    assert (c.k1max, c.k2max, c.n3) == (ref.grid.k1max, ref.grid.k2max, n3)
    assert np.allclose(np.ctypeslib.as_array(c.alpha_c, (n3,)), ref.alpha_c, rtol=1e-15)
    ai = np.ctypeslib.as_array(c.alpha_i, (n3 + 1,))
    assert np.allclose(ai[1:-1], ref.alpha_i, rtol=1e-15)
    sgs = [pd for k, pd in ref.disc if k == "sgs"][0]
    assert (c.sgs_n1, c.sgs_n2) == sgs["dims"]
    adv = [pd for k, pd in ref.disc if k == "adv"][0]
    assert (c.adv_n1, c.adv_n2) == adv["dims"]
    assert np.allclose(np.ctypeslib.as_array(c.sgs_lc, (n3,)), sgs["lc"], rtol=1e-14)
    assert np.allclose(np.ctypeslib.as_array(c.sgs_li, (n3 - 1,)), sgs["li"], rtol=1e-14)
    assert np.allclose(list(c.sgs_factor_strain), sgs["fstrain"], rtol=1e-14)
    assert np.allclose(list(c.sgs_factor_stress), sgs["fstress"], rtol=1e-14)
    assert list(c.sgs_wall) == [1 if w == "rough" else 0 for w in sgs["walls"]]


def test_sgs_rejects_smooth_wall():
    dom = B.Domain((1, 1, 1), B.SmoothWall(), B.SmoothWall())
    with pytest.raises(ValueError, match="rough-wall and free-slip"):
        B.model.Discretization((8, 8, 4), dom, [B.StaticSmagorinskyModel()])


def test_constant_mean_constants_match_oracle():
    mk = lambda M: M.Domain((1.0, 1.0, 2.0), M.SmoothWall(), M.SmoothWall(), M.SinusoidalMapping(0.7))
    d = B.model.Discretization((8, 8, 6), mk(B), B.incompressible_flow(1e-2, constant_flux=(1.5, 0.25)))
    ref = O.Model((8, 8, 6), mk(O), O.incompressible_flow(1e-2, constant_flux=(1.5, 0.25)))
    c = d.config
    means = [pd for k, pd in ref.disc if k == "mean"]
    assert list(c.mean) == [1, 1, 0]
    assert list(c.mean_value)[:2] == [1.5, 0.25]
    assert np.allclose(np.ctypeslib.as_array(c.mean_weight, (6,)), means[0]["weight"], rtol=1e-15)
    assert list(c.source) == [0.0, 0.0, 0.0]
    with pytest.raises(ValueError):
        B.incompressible_flow(1.0, constant_flux=1.0, constant_forcing=1.0)
    with pytest.raises(NotImplementedError):
        B.model.Discretization((8, 8, 6), mk(B), [B.ConstantMean("vel3", 1.0)])


def test_unsupported_processes_rejected_on_host():
    dom = B.Domain((1, 1, 1), B.SmoothWall(), B.SmoothWall())
    with pytest.raises(NotImplementedError):
        B.model.Model.set_method(None, object())
    with pytest.raises(NotImplementedError):
        B.model.Discretization((4, 4, 4), dom, [B.MolecularDiffusion("theta", 1.0)])
    with pytest.raises(ValueError):
        B.model.Discretization((4, 4, 2), dom, [B.Pressure()])


def test_whole_plane_rule_matches_the_lds_layouts():
    """The byte model's replica of the engine's whole-plane eligibility rule:
    the 128x128x64 LES grids (192² advection, 128² SGS) take both whole-plane
    transforms in float32; float64 takes them up to 96² only; non-square and
    odd grids never.  The LDS sizes follow the layouts of transforms_2d.hpp:
    Z rows of max(nk1, ⌈B(A+1)/2⌉) entries spread to ≡ B/2 (mod 16), plus the
    N twiddles (c2r); Z rows of nk1 entries spread the same way, one round of
    the row-pair exchange, the twiddles and the D1 factors (r2c)."""
    W = load_package().workloads
    # c2r, 192² fp32: A = 16, B = 12, zs = 102 -> (192*102 + 192) * 8 B = 158,208 B
    assert W.c2r_fused(192, 192, 64, "fp32") and (192 * 102 + 192) * 8 == 158208
    # r2c, 192² fp32: zs = 70, 32 pairs of 12*17 entries -> 161,792 B <= 160 KiB
    assert W.r2c_fused(192, 192, 64, "fp32") and (192 * 70 + 32 * 12 * 17 + 192 + 64) * 8 == 161792
    assert W.c2r_fused(128, 128, 64, "fp32") and W.r2c_fused(128, 128, 64, "fp32")
    assert not W.c2r_fused(192, 192, 64, "fp64") and not W.r2c_fused(192, 192, 64, "fp64")
    assert W.c2r_fused(96, 96, 33, "fp64")
    assert not W.c2r_fused(18, 15, 7, "fp32") and not W.c2r_fused(30, 30, 11, "fp32")  # non-square / no register FFT
    assert not W.c2r_fused(384, 384, 128, "fp32")  # plane beyond LDS


def _ldlt_chunked_max_index(nk, nz, refine, CL=16):
    """Largest element index of the global loads of the chunked LDLt kernel
    (k_ldlt_chunked, 256-thread blocks of MB systems × NC chunks of CL levels):
    the address expressions of its loads, levels clamped into the column."""
    NC = 1
    while NC * CL < nz:
        NC *= 2
    MB = 256 // NC
    worst = 0
    for t in range(256):
        ty = t // MB
        for blk in range((nk + MB - 1) // MB):
            m = blk * MB + t % MB
            mm = m if m < nk else 0
            for j in range(CL):
                l = ty * CL + j
                lc = min(l, nz - 1)
                idx = [lc * nk + mm, max(lc - 1, 0) * nk + mm, min(lc, max(nz - 2, 0)) * nk + mm]
                if refine:
                    if l > 0:
                        idx.append(min(max(l - 1, 0), max(nz - 2, 0)) * nk + mm)
                    if l + 1 < nz:
                        idx.append(min(l, max(nz - 2, 0)) * nk + mm)
                worst = max(worst, *idx)
    return worst


@pytest.mark.parametrize("nk,nz", [(6, 3), (6, 16), (37, 17), (37, 300), (1024, 64)])
def test_ldlt_chunked_loads_stay_in_bounds(nk, nz):
    """The padding levels of the last chunk (levels nz..NC·CL−1) load at
    clamped levels; the float32 refinement's β below a level once read past
    the end of β for such levels (up to (NC·CL−nz)·nk elements)."""
    for refine in (False, True):
        assert _ldlt_chunked_max_index(nk, nz, refine) < nk * nz
