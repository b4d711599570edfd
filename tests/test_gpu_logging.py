"""StepTimer and ProgressMonitor outputs (src/logging.jl:81-308) on the HIP
engine.  The monitor's samples are pinned by fields whose values are known in
closed form: the midpoint rule of a parabola for the mean velocity, Parseval
for the kinetic energies, the rough-wall law for the wall stress and the
largest grid value of a sine for the advective time scale."""

import io
import json
import math

import numpy as np
import pytest

from conftest import load_package

pytestmark = pytest.mark.gpu

B = load_package()


def sgs_channel(n=8, z0=1e-4):
    dom = B.Domain((2 * math.pi, 2 * math.pi, 1.0), B.RoughWall(z0), B.FreeSlipBoundary())
    procs = B.incompressible_flow(1e-3, sgs_model=B.StaticSmagorinskyModel())
    return B.Model((n, n, n), dom, procs, precision="fp64")


@pytest.mark.parametrize("tstep", [None, 1e-3])
def test_progress_monitor_samples_closed_forms(tstep):
    n, z0, a, w0 = 8, 1e-4, 0.2, 0.7
    h = 1.0 / n
    m = sgs_channel(n, z0)
    m.initialize(vel1=lambda x, y, z: 1 + z * (1 - z) + a * np.cos(x), vel2=0.0,
                 vel3=lambda x, y, z: w0 * np.sin(y))
    out = io.StringIO()
    pm = B.ProgressMonitor(frequency=0, tstep=tstep, out=out)
    pm.start(0.0, 1.0)
    s = dict(pm.check(m, 0.25))
    assert s["Simulation Time"] == (0.25,)
    # midpoint rule of 1 + z(1-z) over n cells: 1 + 1/6 + h^2/12
    mv = s["Mean Velocity"]
    assert mv[0] == pytest.approx(1 + 1 / 6 + h * h / 12, rel=1e-14)
    assert abs(mv[1]) < 1e-15 and abs(mv[2]) < 1e-15
    zc = (np.arange(n) + 0.5) * h
    U = 1 + zc * (1 - zc)
    assert s["Mean KE"][0] == pytest.approx(np.sum(U ** 2), rel=1e-13)
    assert s["Turbulent KE"][0] == pytest.approx(n * a * a / 2, rel=1e-12)
    # rough wall below: factor (U1^2 + a^2/2), factor = -kappa^2 / log(h/2 / z0)^2
    fac = -0.4 ** 2 / math.log(h / 2 / z0) ** 2
    t13, t23 = s["Wall Stress τ₁₃"], s["Wall Stress τ₂₃"]
    assert t13[0] == pytest.approx(fac * (U[0] ** 2 + a * a / 2), rel=1e-12)
    assert t13[1] == 0.0 and abs(t23[0]) < 1e-15 and t23[1] == 0.0
    if tstep is None:
        assert s["Advective Timescale"][0] == pytest.approx(h / w0, rel=1e-13)
    else:
        assert s["Adv. Courant Number"][0] == pytest.approx(tstep * w0 / h, rel=1e-13)
    text = out.getvalue()
    assert "Progress:  25%" in text and "ETA: unknown" in text and "Mean Velocity ╎ 1.17" in text


def test_step_timer_and_monitor_during_evolve(tmp_path):
    dt, nt = 1e-3, 20
    m = sgs_channel()
    ref = sgs_channel()
    for mm in (m, ref):
        mm.initialize(vel1=lambda x, y, z: 1 + z * (1 - z) + 0.1 * np.cos(x) * np.sin(math.pi * z), vel2=0.0,
                      vel3=0.0)
    path = tmp_path / "out" / "timestamps.json"
    out = io.StringIO()
    timer = B.StepTimer(path=str(path), frequency=7)
    pm = B.ProgressMonitor(frequency=0, tstep=dt, out=out, chunk=5)
    t = B.evolve(m, nt * dt, dt, output=[timer, pm])
    B.evolve(ref, nt * dt, dt)
    assert t == pytest.approx(nt * dt, rel=1e-14)
    d = json.loads(path.read_text())
    assert set(d) == {"frequency", "wallTime", "simulationTime"} and d["frequency"] == 7
    # the initial state and every seventh step (logging.jl:108-113)
    np.testing.assert_allclose(d["simulationTime"], np.array([0, 7, 14]) * dt, rtol=1e-13, atol=0)
    assert len(d["wallTime"]) == len(d["simulationTime"]) and np.all(np.diff(d["wallTime"]) >= 0)
    # a report at t1 and after every chunk: at most 5 steps, ending at each stamp
    times = [r["Simulation Time"][0] for r in pm.reports]
    np.testing.assert_allclose(times, np.array([0, 5, 7, 12, 14, 19, 20]) * dt, rtol=1e-13, atol=0)
    assert "100%▕" + "█" * 25 + "▏" in out.getvalue()
    # stepping in chunks leaves the solution as one engine call does
    for f in ("vel1", "vel2", "vel3"):
        np.testing.assert_array_equal(m.state[f], ref.state[f])
