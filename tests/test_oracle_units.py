"""Pins the CPU oracle against the reference's own closed-form tests.

Each test restates one reference test (cited); random draws use NumPy's seeded
generator instead of Julia's RNG, the assertions do not depend on the draws.
"""

import math

import numpy as np
import pytest

import oracle as O
from oracle.model import NODES, advection_terms, set_field
from conftest import isapprox

DS = (3 * 0.5176366701850579, 2 * 0.9360371626346338, 0.10477115957769456)


# -- test/grid_test.jl:87-105 ------------------------------------------------
def test_grid_setup():
    gd = O.Grid((64, 32, 16))
    gd2 = O.Grid((63, 31, 16))
    assert (gd.k1max, gd.k2max, gd.n3c, gd.n3i) == (gd2.k1max, gd2.k2max, gd2.n3c, gd2.n3i)
    assert (gd.k1max, gd.k2max, gd.n3global) == (31, 15, 16)
    assert gd.n3c == 16 and gd.n3i == 15
    k1, k2 = O.wavenumbers(gd)
    assert list(k1) == list(range(32))
    assert list(k2) == list(range(16)) + list(range(-15, 0))
    with pytest.raises(ValueError):
        O.Grid((8, 8, 2))


def test_pdsize_rules():
    # physical_space.jl:42-56 (documented sizes of the configurations)
    for k, n in ((31, 96), (63, 192), (5, 16), (6, 20), (7, 24), (47, 142), (23, 70), (0, 3)):
        g = O.Grid((2 * k + 1, 2 * k + 1, 4))
        assert O.pdsize(g, "quadratic") == (n, n)
    g = O.Grid((128, 128, 4))
    assert O.pdsize(g) == (128, 128)


# -- test/transform_test.jl:1-67 ----------------------------------------------
def _setup(n1, n2, n3):
    gd = O.Grid((n1, n2, n3))
    K1, K2 = gd.fdsize()
    rng = np.random.default_rng(1)
    uh = np.full((K1, K2, gd.n3c), complex(*rng.random(2)))
    uv = np.full((K1, K2, gd.n3i), complex(*rng.random(2)))
    return gd, uh, uv


def test_transform_constant():
    gd, uh, uv = _setup(64, 32, 16)
    dom = O.Domain(DS, None, None)
    n1, n2 = O.pdsize(gd)
    set_field(uh, lambda x, y, z: 1.0, dom, gd, "C", n1, n2)
    set_field(uv, lambda x, y, z: 2.0, dom, gd, "I", n1, n2)
    assert isapprox(uh[0, 0, :], np.ones(gd.n3c))
    assert isapprox(uv[0, 0, :], 2 * np.ones(gd.n3i))
    assert np.abs(uh[1:, 0, :]).max() < 1e-14 and np.abs(uh[:, 1:, :]).max() < 1e-14
    assert np.abs(uv[1:, 0, :]).max() < 1e-14 and np.abs(uv[:, 1:, :]).max() < 1e-14
    assert isapprox(O.get_field(uh, n1, n2), np.ones((n1, n2, gd.n3c)))
    assert isapprox(O.get_field(uv, n1, n2), 2 * np.ones((n1, n2, gd.n3i)))


def test_transform_horizontally_varying():
    gd, uh, uv = _setup(16, 16, 16)
    dom = O.Domain(DS, None, None)
    rng = np.random.default_rng(7)
    k = np.arange(-7, 9)
    ax, ay = rng.random(16), rng.random(16)
    L1, L2 = dom.hsize
    u0 = lambda x, y, z: (np.sum(ax * np.cos(2 * np.pi * k * x / L1)) + np.sum(ay * np.cos(2 * np.pi * k * y / L2)))
    u0n = lambda x, y, z: (np.sum((ax * np.cos(2 * np.pi * k * x / L1))[:-1])
                           + np.sum((ay * np.cos(2 * np.pi * k * y / L2))[:-1]))
    n1, n2 = O.pdsize(gd)
    set_field(uh, u0, dom, gd, "C", n1, n2)
    assert isapprox(uh[0, 0, :], (ax[7] + ay[7]) * np.ones(gd.n3c))
    ref_x = (ax[8:15] + ax[6::-1]) / 2
    ref_y = (ay[8:15] + ay[6::-1]) / 2
    assert isapprox(uh[1:, 0, :], ref_x[:, None] * np.ones((1, gd.n3c)))
    assert isapprox(uh[0, 1:8, :], ref_y[:, None] * np.ones((1, gd.n3c)))
    assert isapprox(uh[0, :7:-1, :], ref_y[:, None] * np.ones((1, gd.n3c)))
    assert isapprox(uh[1:, 1:, :] + 1, np.ones((gd.k1max, 2 * gd.k2max, gd.n3c)))
    pd = O.get_field(uh, n1, n2)
    ref = np.array([[[u0n(x, y, 0.0) for _ in range(pd.shape[2])]
                     for y in np.arange(pd.shape[1]) / pd.shape[1] * L2]
                    for x in np.arange(pd.shape[0]) / pd.shape[0] * L1])
    assert isapprox(pd, ref)


def test_transform_vertically_varying():
    gd, uh, uv = _setup(64, 32, 16)
    dom = O.Domain(DS, None, None)
    n1, n2 = O.pdsize(gd)
    zh = np.linspace(0, DS[2], 2 * gd.n3global + 1)[1:-1:2]
    zv = np.linspace(0, DS[2], 2 * gd.n3global + 1)[2:-2:2]
    set_field(uh, lambda x, y, z: z, dom, gd, "C", n1, n2)
    set_field(uv, lambda x, y, z: z, dom, gd, "I", n1, n2)
    assert isapprox(uh[0, 0, :], zh) and isapprox(uv[0, 0, :], zv)
    assert np.abs(uh[1:]).max() < 1e-14 and np.abs(uv[:, 1:]).max() < 1e-14
    ph, pv = O.get_field(uh, n1, n2), O.get_field(uv, n1, n2)
    assert isapprox(ph[0, 0, :], zh) and isapprox(pv[0, 0, :], zv)
    assert isapprox(ph, ph[:1, :1, :] * np.ones((n1, n2, gd.n3c)))


@pytest.mark.parametrize("n1,n2,K1,K2", [(12, 14, 6, 13), (16, 20, 6, 13), (6, 6, 2, 3), (10, 8, 3, 5), (3, 4, 1, 1)])
def test_transforms_brute_force(n1, n2, K1, K2):
    """get_field/add_field equal the defining DFT sums (physical_space.jl:208-295)."""
    rng = np.random.default_rng(n1 * 100 + n2)
    fd = rng.standard_normal((K1, K2, 2)) + 1j * rng.standard_normal((K1, K2, 2))
    k2max = (K2 - 1) // 2
    k1s = np.arange(K1)
    k2s = np.concatenate([np.arange(k2max + 1), np.arange(-k2max, 0)])
    x1 = np.arange(n1)[:, None, None, None]
    x2 = np.arange(n2)[None, :, None, None]
    # c2r semantics: Hermitian completion in x1, real part of the k1=0 column
    ph = np.exp(2j * np.pi * (k1s[None, None, :, None] * x1 / n1 + k2s[None, None, None, :] * x2 / n2))
    w = np.where(k1s == 0, 1.0, 2.0)[None, None, :, None]
    for L in range(2):
        ref = np.real(np.sum(w * fd[None, None, :, :, L] * ph, axis=(2, 3)))
        assert np.allclose(O.get_field(fd, n1, n2)[:, :, L], ref, atol=1e-12)
    pd = rng.standard_normal((n1, n2, 2))
    acc = np.zeros((K1, K2, 2), complex)
    O.add_field(acc, pd)
    for L in range(2):
        ref = np.sum(pd[:, :, None, None, L] * np.conj(ph), axis=(0, 1)) / (n1 * n2)
        assert np.allclose(acc[:, :, L], ref, atol=1e-12)


# -- test/derivatives_test.jl:156-167 (vorticity) -----------------------------
@pytest.mark.parametrize("n3", [3, 4])
def test_vorticity(n3):
    dom = O.Domain((2 * np.pi, 2 * np.pi, 1.0), O.SmoothWall(), O.SmoothWall())
    m = O.Model((12, 14, n3), dom, [O.MomentumAdvection()])
    m.initialize(vel2=lambda x, y, z: 0.856 * np.sin(2 * x))
    dims = O.pdsize(m.grid, "quadratic")
    _, _, _, o1, o2, o3 = advection_terms(m, m.state, dims)
    x1 = dom.hsize[0] * np.arange(dims[0]) / dims[0]
    ref = 0.856 * 2 * np.cos(2 * x1)[:, None, None] * np.ones((1, dims[1], n3))
    assert isapprox(o3, ref)
    assert np.abs(o1).max() < 1e-12 and np.abs(o2).max() < 1e-12


# -- test/diffusion_test.jl:1-52 ----------------------------------------------
@pytest.mark.parametrize("NZ", [16, 3])
def test_diffusion(NZ):
    rval = 0.9134954544827887
    u0d = lambda x, y, z: np.sin(x) * np.cos(y) * (z * (1 - z)) + rval
    u0n = lambda x, y, z: np.cos(x) + np.sin(y) + (z * (1 - z)) + rval * z
    Lu0d = lambda x, y, z: -(2 * z * (1 - z) + 2) * np.sin(x) * np.cos(y)
    Lu0n = lambda x, y, z: -(np.cos(x) + np.sin(y) + 2)
    nu = 1e-6
    lbc = O.CustomBoundary(vel1=("neumann", 1.0 + rval), vel2=("dirichlet", rval), vel3=("dirichlet", rval))
    ubc = O.CustomBoundary(vel1=("neumann", -1.0 + rval), vel2=("dirichlet", rval), vel3=("dirichlet", rval))
    dom = O.Domain((2 * np.pi, 2 * np.pi, 1.0), lbc, ubc)
    dims = (12, 14, NZ)
    m = O.Model(dims, dom, [O.MolecularDiffusion(v, nu) for v in ("vel1", "vel2", "vel3")])
    rhs = {f: np.full_like(a, 0.3 + 0.1j) for f, a in m.state.items()}
    m.initialize(vel1=u0n, vel2=u0d, vel3=u0d)
    O.compute_rates(m, rhs, m.state)
    xs = (4 / 12 * 2 * np.pi, 6 / 14 * 2 * np.pi)  # isample = (5, 7), 1-based
    x3c, x3i = m.coordinates("vel1", 3), m.coordinates("vel3", 3)
    r1 = O.get_field(rhs["vel1"], 12, 14)
    assert isapprox(r1[4, 6, :], nu * np.array([Lu0n(*xs, z) for z in x3c]))
    r2 = O.get_field(rhs["vel2"], 12, 14)
    assert isapprox(r2[4, 6, :], nu * np.array([Lu0d(*xs, z) for z in x3c]))
    r3 = O.get_field(rhs["vel3"], 12, 14)
    assert isapprox(r3[4, 6, :], nu * np.array([Lu0d(*xs, z) for z in x3i]))


# -- test/pressure_test.jl:1-38 (batched LDLt against a dense solve) -----------
@pytest.mark.parametrize("N", [16, 3])
def test_batch_ldlt(N):
    rng = np.random.default_rng(N)
    nk = 2 * 3
    b = rng.random((nk, N)) + 1j * rng.random((nk, N))
    A = O.BatchLDLt(np.ones((nk, N)), np.zeros((nk, N - 1)))
    assert np.allclose(A.gamma, 1) and np.allclose(A.beta, 0)
    assert np.allclose(A.solve(b.copy()), b)
    dv, ev = rng.random(N), rng.random(N - 1)
    scale = np.arange(1, nk + 1)[:, None]
    A = O.BatchLDLt(scale * dv, scale * ev)
    x = A.solve(b.copy())
    T = np.diag(dv) + np.diag(ev, 1) + np.diag(ev, -1)
    for ik in range(nk):
        assert isapprox((ik + 1) * T @ x[ik], b[ik])


# -- test/pressure_test.jl:40-74 (projection of 1+sin to a constant) ----------
def test_pressure_solver():
    bc3 = 0.5700899056030746
    lbc = O.CustomBoundary(vel1="dirichlet", vel2="dirichlet", vel3=("dirichlet", bc3))
    dom = O.Domain((2 * np.pi, 2 * np.pi, 1.0), lbc, lbc)
    m = O.Model((4, 4, 16), dom, [O.Pressure()])
    m.initialize(vel1=lambda x, y, z: 1 + np.sin(x), vel2=0.0, vel3=0.0)
    O.apply_projections(m, m.state)
    assert isapprox(m["vel1"], np.ones_like(m["vel1"]))
    m.initialize(vel1=0.0, vel2=lambda x, y, z: 1 + np.sin(y), vel3=0.0)
    O.apply_projections(m, m.state)
    assert isapprox(m["vel2"], np.ones_like(m["vel2"]))
    m.initialize(vel1=0.0, vel2=0.0, vel3=0.0)
    O.apply_projections(m, m.state)
    assert isapprox(m["vel3"], bc3 * np.ones_like(m["vel3"]))


def _ke_pressure(n, bc, eta, rng):
    """pressure_test.jl:76-125."""
    mapping = None if eta is None else O.SinusoidalMapping(eta, "symmetric")
    dom = O.Domain((1, 1, 1), bc, bc, mapping)
    m = O.Model((n, n, n), dom, [O.Pressure()])
    rnd = lambda x, y, z: rng.random(np.shape(x))
    m.initialize(vel1=rnd, vel2=rnd, vel3=rnd)
    O.apply_projections(m, m.state)
    before = {f: a.copy() for f, a in m.state.items()}
    m.initialize(vel1=rnd, vel2=rnd, vel3=rnd)
    rates = {f: a.copy() for f, a in m.state.items()}
    O.apply_projections(m, m.state)
    for f in rates:
        rates[f] -= m.state[f]
    n1, n2 = O.pdsize(m.grid)
    e = {f: np.sum(O.get_field(before[f], n1, n2) * O.get_field(rates[f], n1, n2), axis=(0, 1)) for f in rates}
    wc = 1 / O.dx3factors(dom, m.grid, "C")
    wi = 1 / O.dx3factors(dom, m.grid, "I")
    return np.sum(wc * e["vel1"]) + np.sum(wc * e["vel2"]) + np.sum(wi * e["vel3"])


def test_pressure_energy():
    rng = np.random.default_rng(871286)
    for bc in (O.SmoothWall(), O.FreeSlipBoundary()):
        for eta in (None, 0.97):
            assert abs(_ke_pressure(8, bc, eta, rng)) < 1e-11


def test_pressure_bc_errors():
    lbc = O.CustomBoundary(vel1="dirichlet", vel2="dirichlet", vel3="neumann")
    m = O.Model((4, 4, 4), O.Domain((1, 1, 1), lbc, lbc), [O.Pressure()])
    with pytest.raises(ValueError, match="only supports Dirichlet"):
        O.apply_projections(m, m.state)
    lbc = O.CustomBoundary(vel1="dirichlet", vel2="dirichlet", vel3=("dirichlet", 1.0))
    m = O.Model((4, 4, 4), O.Domain((1, 1, 1), lbc, O.SmoothWall()), [O.Pressure()])
    with pytest.raises(ValueError, match="mean flow must be equal"):
        O.apply_projections(m, m.state)


# -- test/les_test.jl:1-48 (closed channel LES, closed forms) ----------------
@pytest.mark.parametrize("n", [16, 3])
def test_closed_channel_les(n):
    z0 = 5e-3
    u1, u2 = 0.64917, 0.18572
    ds = (2 * np.pi, 4 / 3 * np.pi, 2.0)
    dom = O.Domain(ds, O.RoughWall(z0), O.RoughWall(z0))
    m = O.Model((n, n, n), dom, [O.StaticSmagorinskyModel(wall_damping=False)])
    rhs = {f: np.zeros_like(a) for f, a in m.state.items()}
    m.initialize(vel1=u1, vel2=u2)
    O.compute_rates(m, rhs, m.state)
    utot = math.sqrt(u1 ** 2 + u2 ** 2)
    tw = 0.4 ** 2 * utot ** 2 / math.log(ds[2] / n / 2 / z0) ** 2
    prof = np.r_[1, np.zeros(n - 2), 1]
    assert isapprox(rhs["vel1"][0, 0, :], prof * tw * u1 / utot / (-ds[2] / n))
    assert isapprox(rhs["vel2"][0, 0, :], prof * tw * u2 / utot / (-ds[2] / n))

    dudz = 0.40595
    m.initialize(vel1=lambda x, y, z: (z - 1) * dudz)
    O.compute_rates(m, rhs, m.state)
    S13 = 0.5 * dudz
    Stot = math.sqrt(2 * (S13 ** 2 + S13 ** 2))
    gs = np.array([n, n, n]) + (n % 2) * np.array([1, 1, 0])
    Delta = np.cbrt(np.prod(np.array(ds) / gs))
    nuT = (Delta * 0.1) ** 2 * Stot
    sgs = [pd for k, pd in m.disc if k == "sgs"][0]
    assert isapprox(sgs["evc"][0, 0, 1:-1], nuT * np.ones(n - 2))
    assert isapprox(sgs["evi"][0, 0, :], nuT * np.ones(n - 1))
    uw = (1 - ds[2] / n / 2) * dudz
    tw = 0.4 ** 2 * (-uw * uw) / math.log(ds[2] / n / 2 / z0) ** 2
    advw = (2 * nuT * S13 - tw) / (ds[2] / n)
    assert isapprox(rhs["vel1"][0, 0, :], np.r_[advw, np.zeros(n - 2), -advw])
    assert isapprox(rhs["vel2"][0, 0, :] + 1, np.ones(n))
    assert isapprox(rhs["vel3"][0, 0, :] + 1, np.ones(n - 1))
    O.evolve(m, 1e-3 * 10, 1e-3)
    assert all(np.all(np.isfinite(a)) for a in m.state.values())


def test_open_channel_les():
    z0 = 5e-3
    ds = (2 * np.pi, 4 / 3 * np.pi, 1.0)
    dom = O.Domain(ds, O.RoughWall(z0), O.FreeSlipBoundary())
    procs = O.incompressible_flow(1e-5, constant_forcing=(1, 0), sgs_model=O.StaticSmagorinskyModel())
    m = O.Model((16, 16, 16), dom, procs)
    m.initialize(vel1=0.64917, vel2=0.18572)
    O.evolve(m, 1e-3 * 10, 1e-3)
    assert all(np.all(np.isfinite(a)) for a in m.state.values())


def test_divergence_error():
    dom = O.Domain((1, 1, 1), O.SmoothWall(), O.SmoothWall())
    m = O.Model((4, 4, 4), dom, O.incompressible_flow(1.0))
    m.state["vel1"][1, 1, 2] = np.nan
    with pytest.raises(FloatingPointError, match="diverged"):
        O.evolve(m, 1e-3, 1e-3)


# -- golden vectors (tests/golden/make_golden.py) ------------------------------
@pytest.mark.parametrize("name", ["les_12x14x16", "channel_8x8x6", "pressure_4x4x16"])
def test_golden_vectors_reproduce(name):
    """The oracle still produces the committed float64 vectors (regression anchor)."""
    import os
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
    import make_golden as G
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", f"{name}.npz"))
    m, dt = G.build(name)
    for f in m.fields:
        m.state[f][...] = g[f"state_{f}"]
    rates = {f: np.zeros_like(a) for f, a in m.state.items()}
    O.compute_rates(m, rates, m.state)
    for f in m.fields:
        assert np.allclose(rates[f], g[f"rates_{f}"], rtol=0, atol=1e-13 * np.abs(g[f"rates_{f}"]).max())
    O.apply_projections(m, m.state)
    for f in m.fields:
        assert np.allclose(m.state[f], g[f"projected_{f}"], rtol=0, atol=1e-13 * np.abs(g[f"projected_{f}"]).max())
