"""ODE methods of the reference, restated (test infrastructure only).

Restates ``src/ODEMethods.jl`` and ``approxdiv`` from ``src/util.jl:19-24``.
States are lists of NumPy arrays (the reference uses an ``ArrayPartition`` that
aliases the model state, ``src/BoundaryLayerDynamics.jl:203``); every update is
done in place so that the aliasing semantics are kept.
"""

import math

import numpy as np


def approxdiv(a, b):
    """``src/util.jl:19-24``: integer a/b, error if not (approximately) integer."""
    result_float = a / b
    result_int = int(round(float(result_float)))
    if not _isapprox(result_float, type(result_float)(result_int)):
        raise ValueError("Division has a non-integer result")
    return result_int


def _isapprox(x, y):
    # Julia's default isapprox: rtol = sqrt(eps) of the wider operand type, atol = 0
    single = all(isinstance(v, np.float32) for v in (x, y) if isinstance(v, np.floating))
    single = single and any(isinstance(v, np.float32) for v in (x, y))
    eps = np.finfo(np.float32 if single else np.float64).eps
    return x == y or abs(x - y) <= math.sqrt(eps) * max(abs(x), abs(y))


class Euler:  # ODEMethods.jl:21
    pass


class AB2:  # ODEMethods.jl:30
    pass


class SSPRK22:  # ODEMethods.jl:40
    pass


class SSPRK33:  # ODEMethods.jl:50
    pass


def _axpy_list(dst, expr):
    for i in range(len(dst)):
        dst[i][...] = expr(i)


class ODEProblem:
    """``src/ODEMethods.jl:168-185``.  Projects u0 and evaluates du at t1."""

    def __init__(self, rate, u0, tspan, projection=None, checkpoint=False):
        t1, t2 = tspan
        if not t1 <= t2:
            raise ValueError("Integration start time is larger than end time")
        self.rate = rate
        self.projection = projection if projection is not None else (lambda u: u)
        self.u = u0
        self.du = [np.zeros_like(x) for x in u0]
        self.projection(self.u)
        if checkpoint:
            self.rate(self.du, self.u, t1, checkpoint=True)
        else:
            self.rate(self.du, self.u, t1)
        self.t = t1
        self.tmax = t2


def _init_buffer(prob, alg):  # ODEMethods.jl:62-85
    if isinstance(alg, (SSPRK22, SSPRK33)):
        return {"utmp": [np.full_like(x, np.nan) for x in prob.u]}
    if isinstance(alg, AB2):
        return {"duprev": [np.full_like(x, np.nan) for x in prob.du], "last_dt": math.nan}
    return {}


def perform_step(prob, dt, alg, buf):
    """``src/ODEMethods.jl:100-166``."""
    u, du = prob.u, prob.du
    if isinstance(alg, Euler):  # :100-103
        _axpy_list(u, lambda i: u[i] + dt * du[i])
        prob.projection(u)
    elif isinstance(alg, AB2):  # :105-131
        if math.isnan(buf["last_dt"]):
            _axpy_list(u, lambda i: u[i] + dt * du[i])
            prob.projection(u)
            duprev = buf["duprev"]
            _axpy_list(duprev, lambda i: du[i])
            prob.rate(du, u, prob.t + dt)
            _axpy_list(u, lambda i: u[i] + dt * (du[i] - duprev[i]) / 2)
            prob.projection(u)
            buf["last_dt"] = dt
        else:
            if buf["last_dt"] != dt:
                raise ValueError("Cannot change time step for multi-step methods")
            duprev = buf["duprev"]
            _axpy_list(u, lambda i: u[i] + dt / 2 * (3 * du[i] - duprev[i]))
            prob.projection(u)
            _axpy_list(duprev, lambda i: du[i])
    elif isinstance(alg, SSPRK22):  # :133-145
        ut = buf["utmp"]
        _axpy_list(ut, lambda i: u[i] + dt * du[i])
        prob.projection(ut)
        prob.rate(du, ut, prob.t + dt)
        _axpy_list(u, lambda i: (u[i] + ut[i] + dt * du[i]) / 2)
        prob.projection(u)
    elif isinstance(alg, SSPRK33):  # :147-166
        ut = buf["utmp"]
        _axpy_list(ut, lambda i: u[i] + dt * du[i])
        prob.projection(ut)
        prob.rate(du, ut, prob.t + dt)
        _axpy_list(ut, lambda i: (3 * u[i] + ut[i] + dt * du[i]) / 4)
        prob.projection(ut)
        prob.rate(du, ut, prob.t + dt / 2)
        _axpy_list(u, lambda i: (u[i] + 2 * ut[i] + 2 * dt * du[i]) / 3)
        prob.projection(u)
    else:
        raise TypeError(f"unknown ODE method {alg!r}")


def step_to(prob, buf, alg, t, dt, checkpoint):
    """``src/ODEMethods.jl:217-234``."""
    if _isapprox(prob.t, t):
        if checkpoint and prob.t != t:
            raise ValueError(f"Could not hit checkpoint t={t} exactly")
        prob.t = t
        return prob
    nt = approxdiv(t - prob.t, dt)
    for it in range(1, nt + 1):
        perform_step(prob, dt, alg, buf)
        prob.t = t if _isapprox(prob.t + dt, t) else prob.t + dt
        if it == nt and checkpoint:
            prob.rate(prob.du, prob.u, prob.t, checkpoint=True)
        else:
            prob.rate(prob.du, prob.u, prob.t)
    return prob


def solve(prob, alg, dt, checkpoints=None):
    """``src/ODEMethods.jl:198-215``."""
    buf = _init_buffer(prob, alg)
    if checkpoints is None:
        cps = []
    else:
        cps = list(checkpoints)
        if min(cps) <= prob.t or max(cps) > prob.tmax:
            raise ValueError("Checkpoints outside of time range")
        cps = sorted(cps)
    for t in cps:
        step_to(prob, buf, alg, t, dt, True)
    step_to(prob, buf, alg, prob.tmax, dt, False)
    return prob.u
