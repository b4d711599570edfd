"""The reference's laminar analytic tests run through the HIP engine:
test/laminar_test.jl:3-100 with the flow problems of
test/laminar_flow_problems.jl:9-188 (Poiseuille, Couette, Taylor–Green in a
vertical and a horizontal plane, constant flux), in float64, with the
reference's error thresholds and convergence orders.  The same functions run
the oracle in tests/test_oracle_laminar.py."""

import math

import numpy as np
import pytest

from conftest import convergence, load_package
from test_oracle_laminar import couette, nt_viscous, poiseuille

pytestmark = pytest.mark.gpu

B = load_package()


def vel3_at_c_values(d3, lower, upper):
    """I-node values interpolated to C-nodes with Dirichlet boundary values
    (State.interpolate, State.jl:81-99)."""
    ext = np.concatenate([np.full_like(d3[:, :, :1], lower), d3, np.full_like(d3[:, :, :1], upper)], axis=2)
    return (ext[:, :, :-1] + ext[:, :, 1:]) / 2


def laminar_flow_error(Nh, Nv, Nt, u_exact, t, nu, delta, direction, eta, method, vel_bc=0.0, f=0.0):
    """laminar_flow_problems.jl:55-82 on the GPU engine."""
    d = np.array(direction) / np.sqrt(np.sum(np.array(direction) ** 2))
    lbc = B.CustomBoundary(vel1=("dirichlet", -vel_bc * d[0]), vel2=("dirichlet", -vel_bc * d[1]), vel3="dirichlet")
    ubc = B.CustomBoundary(vel1=("dirichlet", vel_bc * d[0]), vel2=("dirichlet", vel_bc * d[1]), vel3="dirichlet")
    mapping = None if eta is None else B.SinusoidalMapping(eta, "symmetric")
    dom = B.Domain((1, 1, 2 * delta), lbc, ubc, mapping)
    m = B.Model((Nh, Nh, Nv), dom, B.incompressible_flow(nu, constant_forcing=tuple(f * d)), precision="fp64")
    B.evolve(m, t, t / Nt, method=method)
    z = m.coordinates("vel1", 3)
    uref = np.array([u_exact(x3, t) for x3 in z])[None, None, :]
    e1 = m["vel1"] - uref * d[0]
    e2 = m["vel2"] - uref * d[1]
    e3 = vel3_at_c_values(m["vel3"], 0.0, 0.0)
    return math.sqrt(np.max(np.abs(e1) ** 2 + np.abs(e2) ** 2 + np.abs(e3) ** 2) / np.max(np.abs(uref) ** 2))


def taylor_green_vortex_error(Nh, Nv, Nt, t=1.0, nu=1.0, a=1.0, b=1.0, g=1.0, U=1.0, W=None, eta=None,
                              method=None):
    """laminar_flow_problems.jl:122-188 on the GPU engine."""
    tg1 = lambda x, y, t: np.sin(x) * np.cos(y) * np.exp(-t)
    tg2 = lambda x, y, t: -np.cos(x) * np.sin(y) * np.exp(-t)
    if W is None:  # vortex in a vertical plane
        s = (a ** 2 + b ** 2 + g ** 2) * nu
        u1 = lambda x, y, z, t: U * a * g * tg1(a * x + b * y, g * z, s * t)
        u2 = lambda x, y, z, t: U * b * g * tg1(a * x + b * y, g * z, s * t)
        u3 = lambda x, y, z, t: U * (a ** 2 + b ** 2) * tg2(a * x + b * y, g * z, s * t)
        bc = B.FreeSlipBoundary()
    else:  # vortex in a horizontal plane
        s = (a ** 2 + b ** 2) * nu
        u1 = lambda x, y, z, t: U * b * tg1(a * x, b * y, s * t)
        u2 = lambda x, y, z, t: U * a * tg2(a * x, b * y, s * t)
        u3 = lambda x, y, z, t: W + 0 * x
        bc = B.CustomBoundary(vel1="neumann", vel2="neumann", vel3=("dirichlet", W))
    ds = (2 * np.pi / a, 2 * np.pi / b, 2 * np.pi / g)
    mapping = None if eta is None else B.SinusoidalMapping(eta, "symmetric")
    dom = B.Domain(ds, bc, bc, mapping)
    m = B.Model((Nh, Nh, Nv), dom, B.incompressible_flow(nu, constant_forcing=0), precision="fp64")
    m.initialize(vel1=lambda x, y, z: u1(x, y, z, 0.0), vel2=lambda x, y, z: u2(x, y, z, 0.0),
                 vel3=lambda x, y, z: u3(x, y, z, 0.0))
    B.evolve(m, t, t / Nt, method=method)
    x1, x2 = m.coordinates("vel1", 1), m.coordinates("vel1", 2)
    zc = m.coordinates("vel1", 3)
    Xc = np.meshgrid(x1, x2, zc, indexing="ij")
    e1 = m["vel1"] - u1(*Xc, t)
    e2 = m["vel2"] - u2(*Xc, t)
    zi = m.coordinates("vel3", 3)
    Xi = np.meshgrid(x1, x2, zi, indexing="ij")
    d3 = m["vel3"] - u3(*Xi, t)
    # the error interpolated with the vel3 boundary value, which the
    # reference requires to be zero (laminar_flow_problems.jl:160,185-190)
    e3 = vel3_at_c_values(d3, 0.0, 0.0)
    uref_max = math.exp(-nu * t * (a ** 2 + b ** 2 + (g ** 2 if W is None else 0)))
    return math.sqrt(np.max(np.abs(e1) ** 2 + np.abs(e2) ** 2 + np.abs(e3) ** 2) / uref_max)


@pytest.mark.parametrize("stretching", [False, True])
def test_laminar_flow_convergence_gpu(stretching):
    """laminar_test.jl:3-65 with the thresholds of :49-62."""
    rng = np.random.default_rng(62275 + int(stretching))
    method = B.SSPRK22()
    Nz_min = 3
    N = [n for n in (11, 14, 18, 24) if n >= Nz_min]
    eta = 0.4 + rng.random() / 5 if stretching else None
    nu = 0.75 + rng.random() / 2
    th = rng.random() * 2 * np.pi
    direction = (math.cos(th), math.sin(th))
    delta = 0.75 + rng.random() / 2
    utau = 0.75 + rng.random() / 2
    t = (delta ** 2 / nu) / 6
    Nt = nt_viscous(N[-1], t=t, delta=delta, nu=nu, eta=eta, Cmax=0.25)
    ep = [laminar_flow_error(3, n, Nt, lambda y, tt: (utau ** 2 * delta / nu) * poiseuille(y / delta, tt * nu / delta ** 2),
                             t, nu, delta, direction, eta, method, f=utau ** 2 / delta) for n in N]
    ec = [laminar_flow_error(3, n, Nt, lambda y, tt: (utau ** 2 * delta / nu) * couette(y / delta, tt * nu / delta ** 2),
                             t, nu, delta, direction, eta, method, vel_bc=utau * utau * delta / nu) for n in N]
    a, b, g, U = (0.75 + rng.random() / 2 for _ in range(4))
    t = 1 / ((a ** 2 + b ** 2 + g ** 2) * nu)
    Nt = nt_viscous(N[-1], delta=np.pi / (2 * b), nu=nu, t=t, Cmax=0.25)
    etgv = [taylor_green_vortex_error(3, n, Nt, t=t, nu=nu, a=a, b=b, g=g, U=U, eta=eta, method=method) for n in N]
    tgh = dict(t=t, nu=nu, a=a, b=b, g=g, U=U, W=0.0, eta=eta)
    eeu = [taylor_green_vortex_error(3, Nz_min, n, method=B.Euler(), **tgh) for n in N]
    eab = [taylor_green_vortex_error(3, Nz_min, n, method=B.AB2(), **tgh) for n in N]
    er2 = [taylor_green_vortex_error(3, Nz_min, n, method=B.SSPRK22(), **tgh) for n in N]
    er3 = [taylor_green_vortex_error(3, Nz_min, n, method=B.SSPRK33(), **tgh) for n in N]
    print("errors", dict(ep=ep[-1], ec=ec[-1], etgv=etgv[-1], eeu=eeu[-1], eab=eab[-1], er2=er2[-1], er3=er3[-1]))
    assert ep[-1] < (1e-3 if stretching else 2e-3)
    assert ec[-1] < (1e-3 if stretching else 5e-4)
    assert etgv[-1] < (7e-3 if stretching else 5e-3)
    convergence(N, ep, order=2, threshold_slope=0.95)
    convergence(N, ec, order=2, threshold_slope=0.9)
    convergence(N, etgv, order=2, threshold_slope=0.94)
    assert eeu[-1] < 2e-2 and eab[-1] < 4e-4 and er2[-1] < 2e-4 and er3[-1] < 2e-6
    convergence(N, eeu, order=1, threshold_slope=0.99)
    convergence(N, eab, order=2, threshold_slope=0.98)
    convergence(N, er2, order=2, threshold_slope=0.99)
    convergence(N, er3, order=3, threshold_slope=0.99)


def test_constant_flux_poiseuille_gpu():
    """laminar_test.jl:67-100: with ConstantMean the channel approaches the
    Poiseuille profile monotonically and faster than with constant forcing."""
    nz = 16
    z = np.linspace(0, 2, 2 * nz + 1)[1::2]
    u_exact = poiseuille(z, math.inf)
    mf = np.sum(u_exact) / nz
    CFL, nu, tend, nit = 0.25, 1.0, 1.0, 8
    dtmax = (2 / nz) ** 2 / nu * CFL
    nt = math.ceil(tend / dtmax / nit)
    dt = tend / (nit * nt)
    dom = B.Domain((4 * np.pi, 2 * np.pi, 2), B.SmoothWall(), B.SmoothWall())
    m1 = B.Model((4, 4, nz), dom, B.incompressible_flow(1.0, constant_forcing=1), precision="fp64")
    m2 = B.Model((4, 4, nz), dom, B.incompressible_flow(1.0, constant_flux=mf), precision="fp64")
    err = lambda m: np.sum(np.abs(m.state["vel1"][0, 0, :] - u_exact)) / nz
    ef, em = [], []
    for _ in range(nit):
        B.evolve(m1, dt * nt, dt)
        ef.append(err(m1))
        B.evolve(m2, dt * nt, dt)
        em.append(err(m2))
    assert all(np.array(em) < np.array(ef))
    assert all(np.diff(em) < 0)
    assert em[-1] < 1e-6
