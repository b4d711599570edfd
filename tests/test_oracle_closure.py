"""Pins the oracle of the learned SGS closure and of its filtered-DNS sampler.

The closure has no reference counterpart; its oracle is pinned by closed forms
(analytic gradients, a shear flow with a constant-output network), by reduction
to the reference's Smagorinsky model (static_smagorinsky.jl:129-249) when the
network output vanishes, and by explicit loops over analytic point values for
the box filter.
"""

import math

import numpy as np

import oracle as O
from oracle import closure as C
from oracle.model import set_field


def _les(n=(12, 10, 8), Cs=0.1, walls=("rough", "freeslip")):
    mk = {"rough": lambda: O.RoughWall(1e-3, 0.4), "freeslip": O.FreeSlipBoundary}
    dom = O.Domain((2 * math.pi, 2 * math.pi, 1.0), mk[walls[0]](), mk[walls[1]]())
    return O.Model(n, dom, O.incompressible_flow(1e-3, constant_forcing=1,
                                                 sgs_model=O.StaticSmagorinskyModel(Cs=Cs, wall_damping=True)))


def _random_state(m, seed=3, amp=1.0):
    rng = np.random.Generator(np.random.PCG64(seed))
    st = {}
    for f, a in m.state.items():
        z = (rng.standard_normal(a.shape) + 1j * rng.standard_normal(a.shape)) * amp
        z[0, :, :] = 0
        km = (a.shape[1] - 1) // 2
        # (k1 = 0, k2 < 0) mirrors (k1 = 0, k2 > 0): the state of a real field
        z[0, 1:km + 1, :] = rng.standard_normal((km, a.shape[2])) + 1j * rng.standard_normal((km, a.shape[2]))
        z[0, km + 1:, :] = np.conj(z[0, 1:km + 1, :][::-1, :])
        z[0, 0, :] = rng.standard_normal(a.shape[2]) + 3.0
        st[f] = z
    return st


def _set(m, fns):
    g = m.grid
    n1, n2 = C.default_size(m)
    st = {}
    for f, fn in fns.items():
        a = np.zeros_like(m.state[f])
        set_field(a, fn, m.domain, g, O.model.NODES[f], n1, n2)
        st[f] = a
    return st


def test_gradient_definition_closed_form():
    """Linear-in-z u, v and quadratic w (w = 0 at both walls): horizontal
    derivatives are exact, the vertical differences are exact for these
    profiles, and the interpolated ∂w/∂x_i differ from the C-node value by the
    known O(h²) term of the two-point mean."""
    m = _les(n=(8, 8, 6), walls=("freeslip", "freeslip"))
    a, b = 0.7, -1.3
    st = _set(m, {"vel1": lambda x, y, z: (1 + a * z) * np.sin(y),
                  "vel2": lambda x, y, z: (2 + b * z) * np.cos(x),
                  "vel3": lambda x, y, z: z * (1 - z) * np.cos(x + y)})
    t = C.physical_terms(m, st)
    g = C.velocity_gradient(m, t)
    n1, n2 = C.default_size(m)
    x = (2 * math.pi * np.arange(n1) / n1)[:, None, None]
    y = (2 * math.pi * np.arange(n2) / n2)[None, :, None]
    h = 1 / 6
    z = (np.arange(6) + 0.5)[None, None, :] * h
    zero = np.zeros((n1, n2, 6))
    np.testing.assert_allclose(g[..., 0], zero, atol=1e-12)                      # ∂u/∂x1
    np.testing.assert_allclose(g[..., 1], (1 + a * z) * np.cos(y) + zero, atol=1e-12)
    np.testing.assert_allclose(g[..., 2], a * np.sin(y) + zero, atol=1e-12)      # ∂u/∂x3 (linear)
    np.testing.assert_allclose(g[..., 3], -(2 + b * z) * np.sin(x) + zero, atol=1e-12)
    np.testing.assert_allclose(g[..., 4], zero, atol=1e-12)
    np.testing.assert_allclose(g[..., 5], b * np.cos(x) + zero, atol=1e-12)
    np.testing.assert_allclose(g[..., 8], (1 - 2 * z) * np.cos(x + y), atol=1e-12)  # exact for quadratics
    # mean of the I-node values of ∂w/∂x1 = −z(1−z) sin(x+y) below/above (0 at the walls)
    zb, za = z - h / 2, z + h / 2
    q = lambda s: s * (1 - s)
    expect = -0.5 * (q(zb) + q(za)) * np.sin(x + y)
    np.testing.assert_allclose(g[..., 6], expect, atol=1e-12)
    assert np.abs(g[..., 6] - (-q(z) * np.sin(x + y))).max() <= h * h / 4 + 1e-12


def test_sampler_against_explicit_loops():
    """Box-filter averages and τ = avg(u_i u_j) − avg(u_i) avg(u_j) from the
    analytic point values, looped explicitly."""
    m = _les(n=(15, 15, 5), walls=("freeslip", "freeslip"))
    fu = lambda x, y, z: 1 + 0.5 * np.sin(x) * np.cos(2 * y) + 0.2 * z
    fv = lambda x, y, z: 0.3 * np.cos(3 * x) - 0.1 * np.sin(y) * z
    fw = lambda x, y, z: z * (1 - z) * (0.5 + np.sin(x + y))
    st = _set(m, {"vel1": fu, "vel2": fv, "vel3": fw})
    F, delta = 4, 0.37
    xw = np.linspace(0.1, 0.5, 5)
    X, Y = C.sample(m, st, F, delta, xw)
    n1, n2 = C.default_size(m)
    g = C.velocity_gradient(m, C.physical_terms(m, st))
    h = 1 / 5
    n1c, n2c = n1 // F, n2 // F
    for l in range(5):
        zc = (l + 0.5) * h
        for i2c in range(n2c):
            for i1c in range(n1c):
                acc = np.zeros(3)
                prod = np.zeros(6)
                G = np.zeros(9)
                for b in range(F):
                    for a in range(F):
                        i1, i2 = F * i1c + a, F * i2c + b
                        x, y = 2 * math.pi * i1 / n1, 2 * math.pi * i2 / n2
                        wc = 0.5 * ((fw(x, y, zc - h / 2) if l > 0 else 0.0) + (fw(x, y, zc + h / 2) if l < 4 else 0.0))
                        u = np.array([fu(x, y, zc), fv(x, y, zc), wc])
                        acc += u
                        prod += [u[0] * u[0], u[1] * u[1], u[2] * u[2], u[0] * u[1], u[0] * u[2], u[1] * u[2]]
                        G += g[i1, i2, l]
                acc /= F * F
                prod /= F * F
                G /= F * F
                tau = prod - np.array([acc[0] ** 2, acc[1] ** 2, acc[2] ** 2, acc[0] * acc[1], acc[0] * acc[2],
                                       acc[1] * acc[2]])
                tau[:3] -= tau[:3].sum() / 3
                sc = delta ** 2 * np.sum(G * G)
                c = i1c + n1c * (i2c + n2c * l)
                np.testing.assert_allclose(Y[c, :6], tau / sc, rtol=1e-9, atol=1e-12)
                np.testing.assert_allclose(Y[c, 15], sc, rtol=1e-12)
                np.testing.assert_allclose(X[c, :9], G / math.sqrt(np.sum(G * G)), rtol=1e-12, atol=1e-14)
                assert X[c, 9] == xw[l]


def test_zero_network_is_smagorinsky_at_vanishing_constant():
    """y ≡ 0 leaves only the rough-wall stress; the reference's Smagorinsky
    model with Cs → 0 gives the same rates (static_smagorinsky.jl:145-200)."""
    m = _les(Cs=0.1)
    st = _random_state(m, amp=0.3)
    got = C.closure_rates(m, st, 0.5, np.full(m.grid.n3c, 0.2), lambda X: np.zeros((X.shape[0], 6)))
    m0 = _les(Cs=1e-14)
    ref = {f: np.zeros_like(st[f]) for f in st}
    O.compute_rates(m0, ref, st)
    for f in st:
        np.testing.assert_allclose(got[f], ref[f], rtol=1e-12, atol=1e-12 * np.abs(ref[f]).max())
    # and the wall stress is really there: without it the first layer differs
    m_fs = _les(Cs=1e-14, walls=("freeslip", "freeslip"))
    ref_fs = {f: np.zeros_like(st[f]) for f in st}
    O.compute_rates(m_fs, ref_fs, st)
    assert np.abs(ref_fs["vel1"][:, :, 0] - got["vel1"][:, :, 0]).max() > 1e-3


def test_constant_network_shear_flow_closed_form():
    """u = U sin(k x2), v = w = 0, network output ≡ c: τ = Δ²U²k²cos²(k x2)·c'
    (c' with the trace of the normal components removed), so the rates are
    vel1 = Δ²U²k³ c12 sin(2k x2), vel2 = Δ²U²k³ c'22 sin(2k x2), vel3 = 0."""
    dom = O.Domain((2 * math.pi, 2 * math.pi, 1.0), O.FreeSlipBoundary(), O.FreeSlipBoundary())
    m = O.Model((9, 9, 6), dom, [O.StaticSmagorinskyModel(Cs=0.1)])
    U, k, delta = 1.7, 1, 0.3
    st = _set(m, {"vel1": lambda x, y, z: U * np.sin(k * y) + 0 * x, "vel2": 0.0, "vel3": 0.0})
    c = np.array([0.3, -0.2, 0.05, 0.11, 0.0, 0.0])
    rates = C.closure_rates(m, st, delta, np.zeros(6), lambda X: np.tile(c, (X.shape[0], 1)))
    cp = c.copy()
    cp[:3] -= c[:3].sum() / 3
    amp = delta ** 2 * U ** 2 * k ** 3
    K1, K2 = m.grid.fdsize()
    e1 = np.zeros((K1, K2, 6), dtype=complex)
    e2 = np.zeros((K1, K2, 6), dtype=complex)
    i2p, i2m = 2 * k, K2 - 2 * k  # k2 = ±2k
    for e, cc in ((e1, cp[3]), (e2, cp[1])):
        e[0, i2p, :] = amp * cc * (-0.5j)
        e[0, i2m, :] = amp * cc * (0.5j)
    np.testing.assert_allclose(rates["vel1"], e1, atol=1e-12)
    np.testing.assert_allclose(rates["vel2"], e2, atol=1e-12)
    np.testing.assert_allclose(rates["vel3"], 0, atol=1e-12)


def test_wall_feature():
    m = _les(n=(8, 8, 4))
    xw = C.wall_feature(m, 0.5)
    z = (np.arange(4) + 0.5) / 4
    np.testing.assert_allclose(xw, 1 / (1 + z / 0.5))
    m2 = _les(n=(8, 8, 4), walls=("freeslip", "freeslip"))
    assert np.all(C.wall_feature(m2, 0.5) == 0)
