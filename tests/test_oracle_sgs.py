"""Pins the Smagorinsky length scale with Mason & Thomson wall damping
(static_smagorinsky.jl:53-87) and the free-slip-top treatment of the strain
and stress (static_smagorinsky.jl:92-106,195-249) with values the mathematics
fixes, for the oracle and for the host mirror's configuration (which computes
the same constants on its own and hands them to the GPU engine).

* κ → ∞ or far from any wall: l = Cs·∛(Δ1Δ2Δ3), the undamped length;
* Cs → ∞: l = κ·d with d the distance to the nearest rough wall (free slip
  contributes no damping);
* one node evaluated by hand for N = 2 and N = 1 in an algebraically different
  form: l = l0·κd/(l0^N + (κd)^N)^(1/N);
* a linear shear u = a·x3 under a free-slip top: |S| at the top C-node is half
  of the interior value (the wall value of S13 is zero), the I-node stress is
  −l²a², and the rate of vel1 vanishes in the interior and equals −α·l²a² at
  the top C-node (zero stress at the free-slip wall).
A flipped exponent, a distance measured from the wrong wall or a mirrored
instead of a zero wall value of S13 fails one of these.
"""

import ctypes
import math

import numpy as np
import pytest

import oracle as O
from conftest import load_package

B = load_package()

L = (2 * math.pi, 2 * math.pi, 1.0)
RES = (8, 8, 6)


def undamped(Cs, res=RES, dims=L):
    """Cs·∛(Δ1Δ2Δ3) on a uniform grid: Δ1 = L1/(2k1max+2), Δ2 likewise,
    Δ3 = L3/n3 (static_smagorinsky.jl:53-59)."""
    k1, k2 = (res[0] - 1) // 2, (res[1] - 1) // 2
    return Cs * ((dims[0] / (2 * k1 + 2)) * (dims[1] / (2 * k2 + 2)) * (dims[2] / res[2])) ** (1 / 3)


def scales(M, Cs, lower, upper, N=2, damping=True, res=RES, dims=L):
    """(lc, li) of the oracle (M = oracle) or of the host mirror's config."""
    dom = M.Domain(dims, lower, upper)
    sgs = M.StaticSmagorinskyModel(Cs=Cs, wall_damping=damping, wall_damping_exponent=N)
    if M is O:
        m = O.Model(res, dom, [sgs])
        pd = [pd for k, pd in m.disc if k == "sgs"][0]
        return np.asarray(pd["lc"]), np.asarray(pd["li"])
    disc = B.model.Discretization(res, dom, [sgs])
    c = disc.config
    n3 = res[2]
    lc = np.ctypeslib.as_array(ctypes.cast(c.sgs_lc, ctypes.POINTER(ctypes.c_double)), shape=(n3,)).copy()
    li = np.ctypeslib.as_array(ctypes.cast(c.sgs_li, ctypes.POINTER(ctypes.c_double)), shape=(n3 - 1,)).copy()
    return lc, li


zc = (np.arange(RES[2]) + 0.5) / RES[2] * L[2]
zi = np.arange(1, RES[2]) / RES[2] * L[2]


@pytest.mark.parametrize("M", [O, B], ids=["oracle", "host"])
def test_damping_vanishes_for_large_kappa(M):
    lc, li = scales(M, 0.1, M.RoughWall(1e-4, 1e12), M.FreeSlipBoundary())
    l0 = undamped(0.1)
    assert np.allclose(lc, l0, rtol=1e-12, atol=0) and np.allclose(li, l0, rtol=1e-12, atol=0)


@pytest.mark.parametrize("M", [O, B], ids=["oracle", "host"])
def test_undamped_without_walls(M):
    lc, li = scales(M, 0.1, M.FreeSlipBoundary(), M.FreeSlipBoundary())
    assert np.allclose(lc, undamped(0.1), rtol=1e-14) and np.allclose(li, undamped(0.1), rtol=1e-14)


@pytest.mark.parametrize("M", [O, B], ids=["oracle", "host"])
def test_large_cs_gives_the_mixing_length(M):
    kappa = 0.37
    lc, li = scales(M, 1e9, M.RoughWall(1e-4, kappa), M.FreeSlipBoundary())
    assert np.allclose(lc, kappa * zc, rtol=1e-12) and np.allclose(li, kappa * zi, rtol=1e-12)
    # two rough walls: the distance to the nearer one
    lc, li = scales(M, 1e9, M.RoughWall(1e-4, kappa), M.RoughWall(1e-4, kappa))
    assert np.allclose(lc, kappa * np.minimum(zc, L[2] - zc), rtol=1e-12)
    assert np.allclose(li, kappa * np.minimum(zi, L[2] - zi), rtol=1e-12)


@pytest.mark.parametrize("M", [O, B], ids=["oracle", "host"])
@pytest.mark.parametrize("N", [1, 2])
def test_damped_node_values(M, N):
    kappa, Cs = 0.4, 0.1
    lc, li = scales(M, Cs, M.RoughWall(1e-4, kappa), M.FreeSlipBoundary(), N=N)
    l0 = undamped(Cs)
    for l_got, d in ((lc[0], zc[0]), (li[2], zi[2]), (lc[-1], zc[-1])):
        lw = kappa * d
        assert math.isclose(l_got, l0 * lw / (l0 ** N + lw ** N) ** (1 / N), rel_tol=1e-13), (N, d)
    # hand value of the first C-node for N = 2: l0 = 0.1·∛(π/4·π/4·1/6),
    # κd = 0.4/12; l = l0·κd/√(l0² + κd²)
    l0h = 0.1 * (math.pi ** 2 / 96) ** (1 / 3)
    kd = 0.4 / 12
    if N == 2:
        assert math.isclose(lc[0], l0h * kd / math.hypot(l0h, kd), rel_tol=1e-13)
    assert lc[0] < min(l0, kappa * zc[0])  # damping only reduces the length


@pytest.mark.parametrize("a", [0.7, -1.3])
def test_linear_shear_under_free_slip_top(a):
    """u = a·x3 (the (0,0) mode of vel1), v = w = 0, SGS process alone, no
    damping: closed forms of |S|, the I-node stress and the vel1 rate."""
    z0, kappa, Cs = 1e-3, 0.4, 0.1
    dom = O.Domain(L, O.RoughWall(z0, kappa), O.FreeSlipBoundary())
    m = O.Model(RES, dom, [O.StaticSmagorinskyModel(Cs=Cs, wall_damping=False)])
    for f in m.fields:
        m.state[f][...] = 0
    m.state["vel1"][0, 0, :] = a * zc
    rates = {f: np.zeros_like(x) for f, x in m.state.items()}
    terms = {}
    O.compute_rates(m, rates, m.state, terms)
    l0 = undamped(Cs)
    evc, evi = terms["evc"], terms["evi"]
    # interior: S13 = a/2 on every I-node, |S| = |a| on I-nodes and interior C-nodes
    assert np.allclose(evi, l0 ** 2 * abs(a), rtol=1e-12)
    assert np.allclose(evc[:, :, 1:-1], l0 ** 2 * abs(a), rtol=1e-12)
    # free-slip top: the wall value of S13 is zero, so |S| is half the interior one
    assert np.allclose(evc[:, :, -1], l0 ** 2 * abs(a) / 2, rtol=1e-12)
    # rough wall: S13 = u1/(2·(Δz/2)·ln((Δz/2)/z0)) at the first C-node, |S| = 2·|S13|
    h = zc[0]
    assert np.allclose(evc[:, :, 0], l0 ** 2 * 2 * abs(a * h / (2 * h * math.log(h / z0))), rtol=1e-12)
    # τ13 on I-nodes = −2·νT·S13 = −l0²·|a|·a; zero at the free-slip top
    assert np.allclose(terms["sgs13"][:, :, 1:-1], -(l0 ** 2) * abs(a) * a, rtol=1e-12)
    assert np.all(terms["sgs13"][:, :, -1] == 0)
    # vel1 rate: −d/dx3 τ13 (α = n3/L3): zero inside, −α·(0 − τ13) at the top C-node
    alpha = RES[2] / L[2]
    r = rates["vel1"][0, 0, :].real
    assert np.allclose(r[1:-1], 0, atol=1e-12 * alpha * l0 ** 2 * a * a)
    assert math.isclose(r[-1], -alpha * (l0 ** 2) * abs(a) * a, rel_tol=1e-12)
    tau_w = -(kappa ** 2) * (a * h) * abs(a * h) / math.log(h / z0) ** 2
    assert math.isclose(r[0], -alpha * (-(l0 ** 2) * abs(a) * a - tau_w), rel_tol=1e-12)
