"""Snapshots and restart through the engine's physical transfers (CBD files,
src/fileio.jl; Snapshots, src/logging.jl:526-602; initialize! from a folder,
src/State.jl:33-66).  The cell-centred output restates
test/output_test.jl:53-91: a band-limited field written with centered = true
must equal the analytic field at the cell centres."""

import math

import numpy as np
import pytest

from conftest import load_package

pytestmark = pytest.mark.gpu

B = load_package()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_shifted_file_output(tmp_path, precision):
    rng = np.random.default_rng(839905175168)
    n1, n2, n3 = 6, 8, 12
    k1max = (n1 - 1) // 2
    C0, Cx, Cy = rng.random(), rng.random((2, k1max)), rng.random((2, k1max))

    def vel0(x, y, z):
        return ((sum(Cx[0, i - 1] * np.sin(i * x + Cx[1, i - 1]) for i in range(1, k1max + 1))
                 + sum(Cy[0, i - 1] * np.sin(i * y + Cy[1, i - 1]) for i in range(1, k1max + 1)) + C0)
                * z * (1 - z))

    dom = B.Domain((2 * math.pi, 2 * math.pi, 1.0), B.FreeSlipBoundary(), B.FreeSlipBoundary())
    m = B.Model((n1, n2, n3), dom, [B.MolecularDiffusion(v, 1.0) for v in ("vel1", "vel2", "vel3")],
                precision=precision)
    m.initialize(vel1=vel0, vel2=vel0, vel3=vel0)
    m.write_snapshot(str(tmp_path / "state-001"), centered=True)
    for f in ("vel1", "vel3"):
        xmin, xmax, x1, x2, x3, data = B.cbd.read_cbd(str(tmp_path / "state-001" / f"{f}.cbd"))
        assert xmin == (0, 0, 0) and np.allclose(xmax, (2 * math.pi, 2 * math.pi, 1.0))
        assert np.allclose(x1, (np.arange(6) + 0.5) * 2 * math.pi / 6)
        want = vel0(x1[:, None, None], x2[None, :, None], x3[None, None, :])
        tol = 1e-12 if precision == "fp64" else 2e-6
        assert np.abs(data - want).max() < tol, (f, np.abs(data - want).max())
        assert data.dtype == (np.float64 if precision == "fp64" else np.float32)


@pytest.mark.parametrize("centered", [True, False])
def test_restart_from_snapshot(tmp_path, centered):
    """A state written as a snapshot and read into a new model is the same
    state (the transforms of band-limited data are exact up to rounding)."""
    W = B.workloads
    n, dom, procs = W.les_neutral(B, (24, 20, 12))
    m = B.Model(n, dom, procs)
    n1, n2 = m.default_size()
    u, v = W.les_initial(m.coordinates("vel1", 3), n1, n2, seed=2)
    m.set_physical("vel1", u)
    m.set_physical("vel2", v)
    m.set_physical("vel3", np.zeros((n1, n2, m.n3 - 1), order="F"))
    B.evolve(m, 10 * 5e-4, 5e-4)
    m.write_snapshot(str(tmp_path / "snap"), centered=centered)
    m2 = B.Model(n, dom, procs)
    m2.initialize_from(str(tmp_path / "snap"))
    a, b = m.state, m2.state
    for f in a:
        err = np.abs(a[f] - b[f]).max() / np.abs(a[f]).max()
        assert err < 1e-12, (f, err)


def test_initialize_from_checks_the_grid(tmp_path):
    dom = B.Domain((1.0, 1.0, 1.0), B.FreeSlipBoundary(), B.FreeSlipBoundary())
    procs = [B.MolecularDiffusion(v, 1.0) for v in ("vel1", "vel2", "vel3")]
    m = B.Model((8, 8, 6), dom, procs)
    m.write_snapshot(str(tmp_path / "s"))
    other = B.Model((8, 8, 7), dom, procs)
    with pytest.raises(ValueError, match="x3"):
        other.initialize_from(str(tmp_path / "s"))
    bigger = B.Model((10, 8, 6), dom, procs)
    with pytest.raises(NotImplementedError, match="default"):
        bigger.initialize_from(str(tmp_path / "s"))


def test_snapshots_output_in_evolve(tmp_path):
    """Snapshots(frequency) writes state-NNN at every multiple of the frequency
    and the last one equals the final state (logging.jl:584-602)."""
    dom = B.Domain((1.0, 1.0, 1.0), B.SmoothWall(), B.SmoothWall())
    m = B.Model((8, 8, 6), dom, B.incompressible_flow(0.1))
    m.initialize(vel1=lambda x, y, z: z * (1 - z) * (1 + 0.1 * np.sin(2 * np.pi * x)))
    out = B.Snapshots(frequency=0.004, path=str(tmp_path / "snaps"), centered=False)
    t = B.evolve(m, 0.012, 1e-3, output=out)
    assert abs(t - 0.012) < 1e-15
    assert sorted(p.name for p in (tmp_path / "snaps").iterdir()) == ["state-001", "state-002", "state-003"]
    _, _, _, _, _, data = B.cbd.read_cbd(str(tmp_path / "snaps" / "state-003" / "vel1.cbd"))
    assert np.allclose(data, m["vel1"], rtol=0, atol=1e-14)


def _les_pair(M, n=(16, 16, 8)):
    dom = M.Domain((2 * math.pi, 2 * math.pi, 1.0), M.RoughWall(1e-4, 0.4), M.FreeSlipBoundary())
    return M.Model(n, dom, M.incompressible_flow(1e-5, constant_forcing=1,
                                                 sgs_model=M.StaticSmagorinskyModel(Cs=0.1, wall_damping=True)))


@pytest.mark.parametrize("precision,rtol", [("fp64", 1e-10), ("fp32", 2e-4)])
def test_mean_profiles_match_oracle(precision, rtol):
    """All 12 profiles integrated on the device over 4 steps against the
    oracle's samples after each step, trapezoidal rule."""
    import oracle as O
    from oracle import profiles as OP
    ref = _les_pair(O)
    gpu = _les_pair(B) if precision == "fp64" else B.Model((16, 16, 8), _les_pair(B).domain,
                                                            _les_pair(B).processes, precision="fp32")
    rng = np.random.default_rng(5)
    W = B.workloads
    n1, n2 = gpu.default_size()
    u, v = W.les_initial(gpu.coordinates("vel1", 3), n1, n2, seed=5)
    gpu.set_physical("vel1", u)
    gpu.set_physical("vel2", v)
    gpu.set_physical("vel3", np.zeros((n1, n2, 7), order="F"))
    gpu.project()
    st = gpu.state
    for f in ref.fields:
        ref.state[f][...] = st[f]
    dt, steps = 5e-4, 4
    names = OP.NAMES
    gpu.init_problem()
    gpu.enable_profiles(names)
    gpu.step(dt, steps)
    got, nsteps = gpu.read_profiles()
    assert nsteps == steps
    samples = [OP.samples(ref, ref.state)]
    for k in range(steps):
        O.evolve(ref, (k * dt, (k + 1) * dt), dt)
        samples.append(OP.samples(ref, ref.state))
    want = OP.trapezoid(samples, dt)
    # profiles that vanish (mean w by continuity) are compared on the scale of the velocity
    floor = rtol * np.abs(want["vel1"]).max()
    bad = {n: np.abs(got[n] - want[n]).max() for n in names
           if np.abs(got[n] - want[n]).max() > max(rtol * np.abs(want[n]).max(), floor)}
    assert not bad, bad


def test_mean_profiles_output_files(tmp_path):
    m = _les_pair(B)
    n1, n2 = m.default_size()
    u, v = B.workloads.les_initial(m.coordinates("vel1", 3), n1, n2, seed=6)
    m.set_physical("vel1", u)
    m.set_physical("vel2", v)
    out = B.MeanProfiles(("vel1", "sgs13"), path=str(tmp_path / "prof" / "profiles"), output_frequency=2e-3)
    B.evolve(m, 4e-3, 5e-4, output=out)
    files = sorted(p.name for p in (tmp_path / "prof").iterdir())
    assert files == ["profiles-01.npz", "profiles-02.npz"]
    d = np.load(str(tmp_path / "prof" / "profiles-02.npz"))
    assert int(d["intervals"]) == 4 and np.allclose(d["timespan"], [2e-3, 4e-3])
    assert d["vel1"].shape == (8,) and d["sgs13"].shape == (9,)
    assert np.all(d["vel1"] > 0) and d["sgs13"][0] < 0 and d["sgs13"][-1] == 0


def test_profiles_need_their_processes():
    dom = B.Domain((1.0, 1.0, 1.0), B.SmoothWall(), B.SmoothWall())
    m = B.Model((8, 8, 6), dom, B.incompressible_flow(0.1))
    m.enable_profiles(("vel1", "adv3"))
    with pytest.raises(B._lib.BldError, match="SGS"):
        m.enable_profiles(("sgs11",))
