"""Pins the oracle's time integration against test/ode_test.jl."""

import math

import numpy as np
import pytest

import oracle as O
from conftest import convergence

METHODS = [(O.Euler, 1), (O.AB2, 2), (O.SSPRK22, 2), (O.SSPRK33, 3)]


def _rate(f):
    def rate(du, u, t, checkpoint=False):
        vals = f(u, t)
        for d, v in zip(du, vals):
            d[...] = v
    return rate


# ODEProblems of ode_test.jl:1-14
PROBLEMS = [
    (_rate(lambda u, t: [1j * u[0]]), (0.0, 2 * np.pi), lambda t: [np.array([np.exp(1j * t)])]),
    (_rate(lambda u, t: [np.sin(t) * np.ones_like(u[0])]), (0.0, 1.0), lambda t: [np.array([2 - np.cos(t)])]),
    (_rate(lambda u, t: [np.array([u[0][1], -2 / t * u[0][1] - u[0][0] ** 5])]), (np.finfo(float).eps, 1.0),
     lambda t: [np.array([1 / np.sqrt(1 + t ** 2 / 3), -np.sqrt(3) * t * (t ** 2 + 3) ** (-1.5)])]),
]


@pytest.mark.parametrize("alg,order", METHODS)
def test_constant_growth(alg, order):
    prob = O.ODEProblem(_rate(lambda u, t: [np.ones(1)]), [np.zeros(1)], (0.0, 1.0))
    O.solve(prob, alg(), 1e-2)
    assert math.isclose(prob.t, 1.0) and np.allclose(prob.u[0], 1.0)


@pytest.mark.parametrize("alg,order", METHODS)
def test_projection(alg, order):
    rate = _rate(lambda u, t: [1 + u[0].imag + 1j])
    proj = lambda u: u[0].__setitem__(Ellipsis, u[0].real)
    prob = O.ODEProblem(rate, [np.zeros(1, complex)], (0.0, 1.0), projection=proj)
    O.solve(prob, alg(), 1 / 8)
    assert np.allclose(prob.u[0], 1.0)


@pytest.mark.parametrize("alg,order", METHODS)
def test_ode_convergence(alg, order):
    N = [8, 16, 32, 64, 128]
    for rate, T, uref in PROBLEMS:
        eps = []
        for nt in N:
            dt = T[-1] / nt
            u0 = [np.array(x, dtype=complex if np.iscomplexobj(x) else float) for x in uref(T[0])]
            prob = O.ODEProblem(rate, u0, T)
            O.solve(prob, alg(), dt)
            eps.append(np.sqrt(np.sum(np.abs(uref(T[-1])[0] - prob.u[0]) ** 2)))
        convergence(N, eps, order=order)


@pytest.mark.parametrize("alg,order", METHODS)
def test_checkpoints(alg, order):
    times = []

    def rate(du, u, t, checkpoint=False):
        if checkpoint:
            times.append(t)
        du[0][...] = 1

    dt = 0.1
    for cps in (None, np.arange(1, 10) * dt, np.arange(1, 3) * 5 * dt, np.arange(1, 7) * 3 * dt):
        times.clear()
        prob = O.ODEProblem(rate, [np.zeros(1)], (0.0, 1.0))
        if cps is None:
            O.solve(prob, alg(), dt)
            assert np.allclose(prob.u[0], 1) and times == []
        elif max(cps) <= 1.0 + 1e-12:
            O.solve(prob, alg(), dt, checkpoints=cps)
            assert np.allclose(prob.u[0], 1) and np.allclose(times, cps)
        else:
            with pytest.raises(ValueError):
                O.solve(prob, alg(), dt, checkpoints=cps)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_rounding(dtype):
    """ode_test.jl:465-477: checkpoints snap to the exact end time."""
    nt = 3
    dt = dtype(1) / nt
    for i in (-5, 5):
        t = dtype(1)
        for _ in range(abs(i)):
            t = np.nextafter(t, dtype(np.inf) if i > 0 else dtype(-np.inf))
        times = []

        def rate(du, u, tt, checkpoint=False):
            if checkpoint:
                times.append(tt)
            du[0][...] = 1

        p = O.ODEProblem(rate, [np.zeros(1, dtype)], (dtype(0), t))
        O.solve(p, O.SSPRK33(), dt, checkpoints=np.linspace(dt, t, nt, dtype=dtype))
        assert np.allclose(p.u[0], 1, rtol=1e-5)
        assert np.allclose(times, [1 / 3, 2 / 3, 1], rtol=1e-5)
        assert times[-1] == t
