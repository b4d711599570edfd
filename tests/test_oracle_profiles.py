"""Pins the oracle of the mean profiles with a linear shear flow u = c·x3 over
a rough wall (free-slip top): every sample has a closed form."""

import math

import numpy as np

import oracle as O
from oracle import profiles as P
from oracle.model import set_field


def test_linear_shear_profiles():
    c, z0, kappa = 1.7, 1e-3, 0.4
    dom = O.Domain((2 * math.pi, 2 * math.pi, 1.0), O.RoughWall(z0, kappa), O.FreeSlipBoundary())
    m = O.Model((8, 8, 6), dom, O.incompressible_flow(1e-3, sgs_model=O.StaticSmagorinskyModel(Cs=0.1,
                                                                                              wall_damping=False)))
    n1, n2 = O.pdsize(m.grid, None)
    set_field(m.state["vel1"], lambda x, y, z: c * z + 0 * x, dom, m.grid, "C", n1, n2)
    s = P.samples(m, m.state)
    zc = (np.arange(6) + 0.5) / 6
    zi = np.arange(1, 6) / 6
    assert np.allclose(s["vel1"], c * zc, atol=1e-14)
    assert np.allclose(s["vel2"], 0, atol=1e-14) and np.allclose(s["vel3"], 0, atol=1e-14)
    assert len(s["vel3"]) == 7 and len(s["adv3"]) == 7 and len(s["sgs13"]) == 7
    # rotation form: adv3 = ω2·(u⁻+u⁺)/2 with ω2 = c, the others vanish
    assert np.allclose(s["adv3"][1:-1], c * c * zi, atol=1e-12)
    assert s["adv3"][0] == 0 and s["adv3"][-1] == 0
    assert np.allclose(s["adv1"], 0, atol=1e-12) and np.allclose(s["adv2"], 0, atol=1e-12)
    # Smagorinsky at the interior I-nodes: S13 = c/2, |S| = c, τ13 = −2 l² |S| S13 = −l² c²
    # with the undamped length l = Cs·∛(Δ1Δ2Δ3) = 0.1·∛(π/4·π/4·1/6)
    li = 0.1 * (math.pi ** 2 / 96) ** (1 / 3)
    assert np.allclose(s["sgs13"][1:-1], -li ** 2 * c * c, rtol=1e-12)
    # wall entry: rough-wall stress −κ²·u·|u|/ln²(z1/z0) with the sign of the lower wall
    z1 = zc[0]
    assert np.isclose(s["sgs13"][0], -(kappa ** 2) * (c * z1) ** 2 / math.log(z1 / z0) ** 2, rtol=1e-12)
    assert s["sgs13"][-1] == 0  # free-slip top
    for n in ("sgs11", "sgs22", "sgs12", "sgs33", "sgs23"):
        assert np.allclose(s[n], 0, atol=1e-12), n


def test_trapezoid():
    a = [{"x": np.array([1.0, 2.0])}, {"x": np.array([3.0, 2.0])}, {"x": np.array([5.0, 0.0])}]
    acc = P.trapezoid(a, 0.5)
    assert np.allclose(acc["x"], [0.25 * 4 + 0.25 * 8, 0.25 * 4 + 0.25 * 2])
