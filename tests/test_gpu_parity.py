"""Parity of the HIP engine (through the C ABI) with the CPU oracle.

Float64 plain form: rates, projections and multi-step runs must agree with the
oracle to 1e-10 relative (the parity bar); single right-hand sides agree
to ~1e-13.  Float32 mixed form: 2e-5 relative for one right-hand side.
"""

import math

import numpy as np
import pytest

import oracle as O
from conftest import load_package

pytestmark = pytest.mark.gpu

B = load_package()

RTOL64 = 1e-10  # float64 tolerance of the step engine against the oracle
RTOL_RHS64 = 1e-12  # one right-hand side / projection
RTOL_RHS32 = 2e-5  # float32 mixed form, one right-hand side


def relerr(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def random_state(model, rng, amp=0.1, decay=0.5, hermitian=True):
    K1, K2 = model.grid.fdsize()
    k1, k2 = O.wavenumbers(model.grid)
    scale = np.exp(-decay * np.sqrt(k1[:, None] ** 2 + k2[None, :] ** 2))[:, :, None]
    out = {}
    for f in model.fields:
        n = model.state[f].shape[2]
        a = (rng.standard_normal((K1, K2, n)) + 1j * rng.standard_normal((K1, K2, n))) * amp * scale
        if hermitian:  # real field: k1 = 0 column Hermitian in k2
            km = model.grid.k2max
            a[0, K2 - km:, :] = np.conj(a[0, 1:km + 1, :][::-1, :])
            a[0, 0, :] = a[0, 0, :].real
        out[f] = np.asfortranarray(a)
    return out


CASES = {
    # closed channel DNS, smooth walls (C1-like, small)
    "channel": dict(res=(12, 10, 8), dims=(2 * np.pi, np.pi, 2.0), lower="SmoothWall", upper="SmoothWall",
                    procs=lambda M: M.incompressible_flow(1 / 180, constant_forcing=1)),
    # closed channel with constant flux (ConstantMean projection, sources.jl:47-77)
    "channel_flux": dict(res=(12, 10, 8), dims=(2 * np.pi, np.pi, 2.0), lower="SmoothWall", upper="SmoothWall",
                         mapping=0.5, procs=lambda M: M.incompressible_flow(1 / 180, constant_flux=(1.0, 0.2))),
    # neutral half channel LES, rough wall + free slip (C2-like, small)
    "les": dict(res=(16, 16, 8), dims=(2 * np.pi, 2 * np.pi, 1.0), lower=("RoughWall", 1e-4), upper="FreeSlipBoundary",
                procs=lambda M: M.incompressible_flow(1e-5, constant_forcing=1,
                                                      sgs_model=M.StaticSmagorinskyModel(Cs=0.1, wall_damping=True))),
    # square LES grids (advection 48x48, SGS 32x32): the fused whole-plane c2r
    # with more than one round of row pairs
    "les_square": dict(res=(32, 32, 8), dims=(2 * np.pi, 2 * np.pi, 1.0), lower=("RoughWall", 1e-4),
                       upper="FreeSlipBoundary",
                       procs=lambda M: M.incompressible_flow(1e-5, constant_forcing=1,
                                                             sgs_model=M.StaticSmagorinskyModel(Cs=0.1))),
    # larger square LES grids: advection 96x96 (fp64: half-plane c2r blocks)
    # and 142x142 (the quadratic rule at k = 47 gives 3k+1 = 142 = 2·71, which
    # has no register FFT: two-pass LDS transforms); the headline's 192² and
    # 128² planes are covered in test_gpu_headline.py
    "les_64": dict(res=(64, 64, 6), dims=(2 * np.pi, 2 * np.pi, 1.0), lower=("RoughWall", 1e-4),
                   upper="FreeSlipBoundary",
                   procs=lambda M: M.incompressible_flow(1e-5, constant_forcing=1,
                                                         sgs_model=M.StaticSmagorinskyModel(Cs=0.1))),
    "les_96": dict(res=(96, 96, 5), dims=(2 * np.pi, 2 * np.pi, 1.0), lower=("RoughWall", 1e-4),
                   upper="FreeSlipBoundary",
                   procs=lambda M: M.incompressible_flow(1e-5, constant_forcing=1,
                                                         sgs_model=M.StaticSmagorinskyModel(Cs=0.1))),
    # closed channel LES with two rough walls, no damping (les_test.jl:1-48)
    "les_closed": dict(res=(12, 14, 6), dims=(2 * np.pi, 4 / 3 * np.pi, 2.0), lower=("RoughWall", 5e-3),
                       upper=("RoughWall", 5e-3),
                       procs=lambda M: [M.StaticSmagorinskyModel(wall_damping=False)]),
    # SGS with quadratic dealiasing (both spaces at the same size)
    "les_dealiased": dict(res=(10, 12, 5), dims=(2.0, 3.0, 1.0), lower=("RoughWall", 1e-3), upper="FreeSlipBoundary",
                          procs=lambda M: M.incompressible_flow(1e-3, constant_forcing=(1, 0.5),
                                                                sgs_model=M.StaticSmagorinskyModel(dealiasing="quadratic"))),
    # stretched grid, free slip both sides
    "stretched": dict(res=(8, 12, 10), dims=(1.0, 2.0, 1.0), lower="SmoothWall", upper="SmoothWall",
                      mapping=0.7, procs=lambda M: M.incompressible_flow(0.05, constant_forcing=2)),
    # the fewest vertical layers the grid allows (Grids.jl:28): every boundary
    # stencil reaches across the whole column
    "les_min": dict(res=(8, 8, 3), dims=(2 * np.pi, 2 * np.pi, 1.0), lower=("RoughWall", 1e-4),
                    upper="FreeSlipBoundary",
                    procs=lambda M: M.incompressible_flow(1e-5, constant_forcing=1,
                                                          sgs_model=M.StaticSmagorinskyModel(Cs=0.1, wall_damping=True))),
    "channel_min": dict(res=(6, 8, 3), dims=(2 * np.pi, np.pi, 2.0), lower="SmoothWall", upper="SmoothWall",
                        procs=lambda M: M.incompressible_flow(1 / 180, constant_forcing=1)),
    # odd physical sizes (k = 1 → quadratic size 6; k = 2 → 8) and a single resolved mode
    "tiny": dict(res=(3, 5, 4), dims=(1.0, 1.0, 1.0), lower="SmoothWall", upper="FreeSlipBoundary",
                 procs=lambda M: M.incompressible_flow(0.1)),
    # custom boundaries: Neumann/Dirichlet values in diffusion, no advection
    "custom": dict(res=(12, 14, 7), dims=(2 * np.pi, 2 * np.pi, 1.0),
                   lower=("Custom", dict(vel1=("neumann", 1.3), vel2=("dirichlet", 0.4), vel3=("dirichlet", 0.2))),
                   upper=("Custom", dict(vel1=("neumann", -0.7), vel2=("dirichlet", 0.9), vel3=("dirichlet", 0.2))),
                   procs=lambda M: [M.MolecularDiffusion(v, 0.3) for v in ("vel1", "vel2", "vel3")]
                   + [M.Pressure(), M.ConstantSource("vel2", 0.25)]),
}


def _boundary(M, spec):
    if isinstance(spec, str):
        return getattr(M, spec)()
    if spec[0] == "Custom":
        return M.CustomBoundary(**spec[1])
    return getattr(M, spec[0])(spec[1])


def make(M, name, **kw):
    c = CASES[name]
    mapping = M.SinusoidalMapping(c["mapping"]) if "mapping" in c else None
    dom = M.Domain(c["dims"], _boundary(M, c["lower"]), _boundary(M, c["upper"]), mapping)
    return M.Model(c["res"], dom, c["procs"](M), **kw)


def pair(name, precision="fp64", seed=1, **kw):
    ref = make(O, name)
    gpu = make(B, name, precision=precision)
    st = random_state(ref, np.random.default_rng(seed), **kw)
    for f in ref.fields:
        ref.state[f][...] = st[f]
    gpu.set_state(**st)
    return ref, gpu


@pytest.mark.parametrize("name", list(CASES))
def test_state_roundtrip(name):
    ref, gpu = pair(name)
    got = gpu.state
    for f in ref.fields:
        assert np.array_equal(got[f], ref.state[f])


@pytest.mark.parametrize("name", list(CASES))
def test_rates_fp64(name):
    ref, gpu = pair(name, seed=3)
    rates = {f: np.zeros_like(a) for f, a in ref.state.items()}
    O.compute_rates(ref, rates, ref.state)
    got = gpu.rates()
    for f in ref.fields:
        assert relerr(got[f], rates[f]) < RTOL_RHS64, (name, f, relerr(got[f], rates[f]))


@pytest.mark.parametrize("name", ["channel", "les", "custom", "stretched", "les_square"])
def test_rates_nonhermitian_input(name):
    """c2r semantics for a k1=0 column that is not Hermitian (FFTW/pocketfft
    take the real part, physical_space.jl:293)."""
    ref, gpu = pair(name, seed=4, hermitian=False)
    rates = {f: np.zeros_like(a) for f, a in ref.state.items()}
    O.compute_rates(ref, rates, ref.state)
    got = gpu.rates()
    for f in ref.fields:
        assert relerr(got[f], rates[f]) < RTOL_RHS64, (name, f)


@pytest.mark.parametrize("name", ["channel", "les", "custom", "stretched", "tiny", "channel_flux", "les_min",
                                  "channel_min"])
def test_projection_fp64(name):
    ref, gpu = pair(name, seed=5)
    O.apply_projections(ref, ref.state)
    gpu.project()
    got = gpu.state
    for f in ref.fields:
        assert relerr(got[f], ref.state[f]) < RTOL_RHS64, (name, f)


@pytest.mark.parametrize("name,steps,dt", [("channel", 5, 1e-3), ("les", 5, 5e-4), ("stretched", 4, 1e-3),
                                           ("channel_flux", 5, 1e-3),
                                           ("les_dealiased", 3, 1e-3), ("tiny", 6, 1e-2), ("custom", 4, 1e-3),
                                           ("les_min", 4, 5e-4), ("channel_min", 4, 1e-3)])
@pytest.mark.parametrize("graph", [False, True])
def test_evolve_fp64(name, steps, dt, graph):
    ref, gpu = pair(name, seed=6)
    O.evolve(ref, steps * dt, dt)
    B.evolve(gpu, steps * dt, dt, graph=graph)
    got = gpu.state
    for f in ref.fields:
        assert relerr(got[f], ref.state[f]) < RTOL64, (name, f, relerr(got[f], ref.state[f]))


@pytest.mark.parametrize("n3,precision", [(100, "fp64"), (300, "fp64"), (1100, "fp64"), (300, "fp32"),
                                          (1100, "fp32")])
def test_tall_columns(n3, precision):
    """Columns past the 64 levels of the default stage chunking: chunks of 8
    and 16 levels (the float64 (0,0) column pass of the mixed form in chunks
    of 2 and 4), and beyond 1024 levels the sequential stage kernel; a closed
    channel, 2 SSPRK33 steps against the oracle."""
    def build(M, **kw):
        dom = M.Domain((2 * np.pi, np.pi, 2.0), M.SmoothWall(), M.SmoothWall())
        return M.Model((4, 6, n3), dom, M.incompressible_flow(1 / 180, constant_forcing=1), **kw)
    ref, gpu = build(O), build(B, precision=precision)
    st = random_state(ref, np.random.default_rng(n3))
    for f in ref.fields:
        ref.state[f][...] = st[f]
    gpu.set_state(**st)
    dt = 1e-4
    O.evolve(ref, 2 * dt, dt)
    B.evolve(gpu, 2 * dt, dt)
    got = gpu.state
    tol = RTOL64 if precision == "fp64" else 1e-5
    for f in ref.fields:
        assert relerr(got[f], ref.state[f]) < tol, (f, relerr(got[f], ref.state[f]))


@pytest.mark.parametrize("name", ["channel", "les", "tiny"])
def test_physical_transfer(name):
    ref, gpu = pair(name, seed=7)
    n1, n2 = O.pdsize(ref.grid)
    for f in ref.fields:
        got = gpu[f]
        exp = O.get_field(ref.state[f], n1, n2)
        assert relerr(got, exp) < 1e-13
    rng = np.random.default_rng(8)
    for f in ref.fields:
        vals = rng.standard_normal((n1, n2, ref.state[f].shape[2]))
        gpu.set_physical(f, vals)
        exp = np.zeros_like(ref.state[f])
        O.add_field(exp, vals)
        assert relerr(gpu.state[f], exp) < 1e-13


@pytest.mark.parametrize("name", ["channel", "les", "custom", "les_square", "les_64", "les_96", "les_min"])
def test_rates_fp32_mixed(name):
    ref, gpu = pair(name, precision="fp32", seed=9)
    rates = {f: np.zeros_like(a) for f, a in ref.state.items()}
    O.compute_rates(ref, rates, ref.state)
    got = gpu.rates()
    for f in ref.fields:
        assert relerr(got[f], rates[f]) < RTOL_RHS32, (name, f, relerr(got[f], rates[f]))


def test_evolve_fp32_mixed_mean_mode():
    """The (0,0) mode is advanced in float64 in the mixed form: the mean
    profile of a forced channel keeps float64 accuracy where the fluctuations
    carry float32 rounding."""
    ref, gpu = pair("channel", precision="fp32", seed=10, amp=0.02)
    ref.state["vel1"][0, 0, :] += 1.5
    st = gpu.state
    st["vel1"][0, 0, :] += 1.5
    gpu.set_state(**st)
    O.evolve(ref, 10 * 1e-3, 1e-3)
    B.evolve(gpu, 10 * 1e-3, 1e-3)
    got = gpu.state
    mean_err = relerr(got["vel1"][0, 0, :], ref.state["vel1"][0, 0, :])
    assert mean_err < 1e-9, mean_err
    for f in ref.fields:
        assert relerr(got[f], ref.state[f]) < 1e-4


def test_divergence_detected():
    ref, gpu = pair("channel", seed=11)
    st = gpu.state
    st["vel1"][1, 1, 2] = np.nan
    gpu.set_state(**st)
    with pytest.raises(B._lib.BldError, match="diverged"):
        gpu.rates()


def oracle_divergence_time(ref, dt, nmax):
    """Time of the oracle's problem when its rate! raises "diverged"
    (ODEMethods.jl:217-231, BoundaryLayerDynamics.jl:155-157), or None."""
    import oracle.model as OM
    u0 = [ref.state[f] for f in ref.fields]
    prob = O.ODEProblem(OM.rate_fn(ref), u0, (0.0, nmax * dt), projection=OM.projection_fn(ref), checkpoint=True)
    try:
        with np.errstate(all="ignore"):
            O.solve(prob, O.SSPRK33(), dt)
    except FloatingPointError:
        return prob.t
    return None


@pytest.mark.parametrize("graph", [False, True])
def test_divergence_reported_at_the_reference_step(graph):
    """A state far beyond the advective stability limit overflows within a
    few steps: the step call stops there with "diverged" and the time the
    reference's problem has reached when it raises (every rate is checked by
    the kernel that consumes it, inside the captured graph too)."""
    ref, gpu = pair("channel", seed=11, amp=3.0)
    dt, n = 0.05, 100
    t_ref = oracle_divergence_time(ref, dt, n)
    assert t_ref is not None and t_ref < 10 * dt
    gpu.init_problem()
    with pytest.raises(B._lib.BldError, match="The simulation has diverged.") as ei:
        gpu.step(dt, n, t=0.0, t_end=n * dt, graph=graph)
    assert abs(ei.value.t - t_ref) < 1e-9 * dt, (ei.value.t, t_ref)


def test_divergence_after_several_step_chunks():
    """A slower blow-up (well over 32 steps, the length of the step chunks
    after which the flag is read): graph replays and single launches stop at
    the same step, near the oracle's (the trajectories separate by rounding
    before the overflow), and the timing hook reports the divergence too."""
    dt, n = 0.05, 400
    ts = []
    for graph in (False, True):
        ref, gpu = pair("channel", seed=11, amp=1.0)
        gpu.init_problem()
        with pytest.raises(B._lib.BldError, match="diverged") as ei:
            gpu.step(dt, n, t=0.0, t_end=n * dt, graph=graph)
        ts.append(ei.value.t)
    assert ts[0] == ts[1], ts
    assert 40 * dt < ts[0] < n * dt, ts
    ref, _ = pair("channel", seed=11, amp=1.0)
    t_ref = oracle_divergence_time(ref, dt, n)
    assert t_ref is not None and abs(ts[0] - t_ref) <= 0.2 * t_ref, (ts, t_ref)
    _, gpu = pair("channel", seed=11, amp=1.0)
    gpu.init_problem()
    with pytest.raises(B._lib.BldError, match="diverged"):
        gpu.time_steps(dt, n, graph=True)


def test_unsupported_configs():
    lbc = B.CustomBoundary(vel1="dirichlet", vel2="dirichlet", vel3=("dirichlet", 0.5))
    with pytest.raises(B._lib.BldError, match="non-zero w"):
        B.Model((4, 4, 4), B.Domain((1, 1, 1), lbc, lbc), [B.MomentumAdvection()])
    with pytest.raises(B._lib.BldError, match="mean flow must be equal"):
        B.Model((4, 4, 4), B.Domain((1, 1, 1), lbc, B.SmoothWall()), [B.Pressure()])
    nbc = B.CustomBoundary(vel1="dirichlet", vel2="dirichlet", vel3="neumann")
    with pytest.raises(B._lib.BldError, match="only supports Dirichlet"):
        B.Model((4, 4, 4), B.Domain((1, 1, 1), nbc, nbc), [B.Pressure()])


# -- batched LDLt (pressure.jl:38-201, pressure_test.jl:1-38) --------------------
@pytest.mark.parametrize("N", [16, 3, 64])
def test_ldlt_dense(N):
    rng = np.random.default_rng(N)
    nk = 6
    dv, ev = rng.random(N) + 2.5, rng.random(N - 1)
    scale = np.arange(1, nk + 1)[:, None]
    b = rng.random((nk, N)) + 1j * rng.random((nk, N))
    A = B.BatchLDLt(scale * dv, scale * ev)
    x = A.solve(b)
    T = np.diag(dv) + np.diag(ev, 1) + np.diag(ev, -1)
    for ik in range(nk):
        assert relerr((ik + 1) * T @ x[ik], b[ik]) < 1e-13
    xo = O.BatchLDLt(scale * dv, scale * ev).solve(b.copy())
    assert relerr(x, xo) < 1e-14


@pytest.mark.parametrize("nk,nz", [(10, 1), (1, 12)])
def test_ldlt_solve_keeps_its_input(nk, nz):
    """One level or one system: the transposed right-hand side is contiguous
    already, and the in-place device solve must still not overwrite it."""
    rng = np.random.default_rng(nk + nz)
    dv, ev = -(rng.random((nk, nz)) + 2.5), rng.random((nk, nz - 1))
    b = rng.standard_normal((nk, nz)) + 1j * rng.standard_normal((nk, nz))
    b0 = b.copy()
    x = B.BatchLDLt(dv, ev).solve(b)
    assert np.array_equal(b, b0)
    assert relerr(x, O.BatchLDLt(dv, ev).solve(b0.copy())) < 1e-14


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 1e-6)])
@pytest.mark.parametrize("nz", [17, 300])
def test_ldlt_tall_systems(precision, tol, nz):
    """Systems taller than one chunk of levels and not a multiple of it: the
    chunked solve with its combine over chunks (and, in float32, the
    refinement step across chunk edges) against the oracle's sequential
    elimination (pressure.jl:106-201)."""
    rng = np.random.default_rng(nz)
    nk = 37
    dv = -(rng.random((nk, nz)) + 2.5)
    ev = rng.random((nk, nz - 1))
    b = rng.random((nk, nz)) + 1j * rng.random((nk, nz))
    x = B.BatchLDLt(dv, ev, precision=precision).solve(b)
    xo = O.BatchLDLt(dv, ev).solve(b.copy())
    assert relerr(x, xo) < tol


def pressure_operator(nk1, nk2, nz, L=(2 * np.pi, 2 * np.pi), L3=1.0):
    """The pressure matrix of pressure.jl:231-296 for a uniform grid."""
    K1, K2 = nk1, nk2
    k1 = np.arange(K1)
    k2 = np.concatenate([np.arange((K2 + 1) // 2), np.arange(-(K2 // 2), 0)])
    a = nz / L3
    d3g3 = np.full(nz, -2 * a)
    d3g3[0] = d3g3[-1] = -a
    kk1, kk2 = np.tile(k1, K2), np.repeat(k2, K1)
    shift = 4 * np.pi ** 2 * (kk1 ** 2 / L[0] ** 2 + kk2 ** 2 / L[1] ** 2)
    dvs = d3g3[None, :] - shift[:, None] / a
    dvs[0] = d3g3
    dvs[0, -1] = -3 * a
    evs = np.full((K1 * K2, nz - 1), a)
    return dvs, evs


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 1e-6)])
def test_ldlt_pressure_operator(precision, tol):
    """Pressure systems (C3 shape at reduced size), float64 and float32 + one
    refinement step, against the oracle's float64 solution."""
    dvs, evs = pressure_operator(32, 32, 64)
    rng = np.random.default_rng(12)
    b = rng.standard_normal(dvs.shape) + 1j * rng.standard_normal(dvs.shape)
    b[0, -1] = 0
    x = B.BatchLDLt(dvs, evs, precision=precision).solve(b)
    xo = O.BatchLDLt(dvs, evs).solve(b.copy())
    err = np.abs(x - xo).max() / np.abs(xo).max()
    assert err < tol, err


# -- committed golden vectors (tests/golden/make_golden.py) ------------------------
@pytest.mark.parametrize("name", ["les_12x14x16", "channel_8x8x6", "pressure_4x4x16"])
def test_golden_vectors(name):
    import os
    import sys
    here = os.path.join(os.path.dirname(__file__), "golden")
    sys.path.insert(0, here)
    import make_golden as G
    g = np.load(os.path.join(here, f"{name}.npz"))
    ref, dt = G.build(name)
    res, dims, lo, hi, procs, _ = G.SPECS[name]
    bnd = lambda s: B.RoughWall(s[1]) if s[0] == "rough" else (B.FreeSlipBoundary() if s[0] == "freeslip"
                                                               else B.SmoothWall())
    # the same processes through the host mirror
    if name.startswith("les"):
        pr = B.incompressible_flow(1e-4, constant_forcing=(1, 0.2), sgs_model=B.StaticSmagorinskyModel())
    elif name.startswith("channel"):
        pr = B.incompressible_flow(1 / 180, constant_forcing=1)
    else:
        pr = [B.Pressure()]
    gpu = B.Model(res, B.Domain(dims, bnd(lo), bnd(hi)), pr)
    st = {f: g[f"state_{f}"] for f in gpu.fields}
    gpu.set_state(**st)
    got = gpu.rates()
    for f in gpu.fields:
        assert relerr(got[f], g[f"rates_{f}"]) < RTOL_RHS64
    gpu.project()
    got = gpu.state
    for f in gpu.fields:
        assert relerr(got[f], g[f"projected_{f}"]) < RTOL_RHS64
    if dt > 0:
        gpu.set_state(**st)
        B.evolve(gpu, dt, dt)
        got = gpu.state
        for f in gpu.fields:
            assert relerr(got[f], g[f"step_{f}"]) < RTOL64


# -- ensembles: several models of uneven size advanced by the same launches ----------
def test_ensemble_matches_oracle():
    names = ["channel", "les", "stretched", "tiny"]
    refs, specs, states = [], [], []
    for i, name in enumerate(names):
        ref = make(O, name)
        st = random_state(ref, np.random.default_rng(40 + i))
        for f in ref.fields:
            ref.state[f][...] = st[f]
        refs.append(ref)
        states.append(st)
        c = CASES[name]
        mapping = B.SinusoidalMapping(c["mapping"]) if "mapping" in c else None
        dom = B.Domain(c["dims"], _boundary(B, c["lower"]), _boundary(B, c["upper"]), mapping)
        specs.append((c["res"], dom, c["procs"](B)))
    ens = B.Ensemble(specs, precision="fp64")
    assert len(ens) == len(names)
    for m, st in zip(ens.members, states):
        m.set_state(**st)
    dt, steps = 1e-3, 3
    ens.init_problem()
    ens.step(dt, steps)
    for ref, m, name in zip(refs, ens.members, names):
        O.evolve(ref, steps * dt, dt)
        got = m.state
        for f in ref.fields:
            assert relerr(got[f], ref.state[f]) < RTOL64, (name, f, relerr(got[f], ref.state[f]))


def test_ensemble_fp32_members_match_oracle():
    """Mixed-form ensemble members (uneven LES grids: whole-plane transforms
    of several lengths in shared launches, float64 (0,0) columns and pressure)
    against the oracle's float64 steps, within the float32 bound of the
    headline tests."""
    names = ["les", "les_square", "les_closed"]
    refs, specs, states = [], [], []
    for i, name in enumerate(names):
        ref = make(O, name)
        st = random_state(ref, np.random.default_rng(70 + i))
        for f in ref.fields:
            ref.state[f][...] = st[f]
        refs.append(ref)
        states.append(st)
        c = CASES[name]
        dom = B.Domain(c["dims"], _boundary(B, c["lower"]), _boundary(B, c["upper"]))
        specs.append((c["res"], dom, c["procs"](B)))
    ens = B.Ensemble(specs, precision="fp32")
    for m, st in zip(ens.members, states):
        m.set_state(**st)
    dt, steps = 5e-4, 3
    ens.init_problem()
    ens.step(dt, steps)
    for ref, m, name in zip(refs, ens.members, names):
        O.evolve(ref, steps * dt, dt)
        got = m.state
        for f in ref.fields:
            assert relerr(got[f], ref.state[f]) < 1e-5, (name, f, relerr(got[f], ref.state[f]))


def test_ensemble_fp32_matches_single_models():
    """Uneven LES grids in one fp32 ensemble equal the same models run alone."""
    res_list = [(12, 12, 6), (16, 16, 8), (20, 20, 10)]
    W = B.workloads
    specs, singles, states = [], [], []
    for i, res in enumerate(res_list):
        n, dom, procs = W.les_neutral(B, res, Cs=0.08 + 0.03 * i)
        specs.append((n, dom, procs))
        single = B.Model(n, dom, procs, precision="fp32")
        ref = make(O, "les")  # only for shapes
        n1, n2 = single.default_size()
        u, v = W.les_initial(single.coordinates("vel1", 3), n1, n2, seed=50 + i)
        single.set_physical("vel1", u)
        single.set_physical("vel2", v)
        single.set_physical("vel3", np.zeros((n1, n2, single.n3 - 1), order="F"))
        states.append(single.state)
        singles.append(single)
    ens = B.Ensemble(specs, precision="fp32")
    for m, st in zip(ens.members, states):
        m.set_state(**st)
    ens.init_problem()
    ens.step(5e-4, 4)
    for s in singles:
        s.init_problem()
        s.step(5e-4, 4)
    for m, s in zip(ens.members, singles):
        a, b = m.state, s.state
        for f in m.fields:
            assert relerr(a[f], b[f]) < 1e-5, (f, relerr(a[f], b[f]))


@pytest.mark.parametrize("name", ["les", "channel", "les_square", "les_64"])
def test_autotune_keeps_results(name):
    ref, gpu = pair(name, seed=60)
    before = gpu.rates()
    gpu.autotune(reps=2)
    tun = gpu.tuning()
    assert all(len(v) >= 1 for v in tun.values())
    after = gpu.rates()
    for f in ref.fields:
        assert relerr(after[f], before[f]) < 1e-14
    O.evolve(ref, 3e-3, 1e-3)
    B.evolve(gpu, 3e-3, 1e-3)
    got = gpu.state
    for f in ref.fields:
        assert relerr(got[f], ref.state[f]) < RTOL64


@pytest.mark.parametrize("name", ["les_square", "channel"])
def test_kernel_replay_keeps_rates(name):
    """bld_time_kernel replays a kernel class on its own (measurement hook):
    positive durations, the dispatch count of one right-hand side, and the
    rates unchanged afterwards (the accumulating r2c classes are recomputed)."""
    ref, gpu = pair(name, precision="fp32", seed=61)
    before = gpu.rates()
    for k in ("c2r_2d", "r2c_2d", "c2r_x2", "r2c_x2", "linear", "adv_products", "sgs_viscosity", "sgs_stress"):
        us, n = gpu.time_kernel(k, reps=3)
        assert n >= 0 and (us > 0) == (n > 0), (k, us, n)
    with pytest.raises(B._lib.BldError):
        gpu.time_kernel("stage_project")
    after = gpu.rates()
    for f in ref.fields:
        assert relerr(after[f], before[f]) < 1e-6


@pytest.mark.parametrize("method", ["Euler", "SSPRK22", "AB2"])
@pytest.mark.parametrize("name,dt", [("channel", 1e-3), ("les", 5e-4), ("channel_flux", 1e-3)])
@pytest.mark.parametrize("graph", [False, True])
def test_other_methods_fp64(method, name, dt, graph):
    """perform_step! of Euler, SSPRK22 and AB2 (ODEMethods.jl:100-145) against
    the oracle's, 5 steps (AB2: the Runge-Kutta start and 4 regular steps)."""
    ref, gpu = pair(name, seed=9)
    O.evolve(ref, 5 * dt, dt, method=getattr(O, method)())
    B.evolve(gpu, 5 * dt, dt, method=getattr(B, method)(), graph=graph)
    got = gpu.state
    for f in ref.fields:
        assert relerr(got[f], ref.state[f]) < RTOL64, (method, name, f, relerr(got[f], ref.state[f]))


@pytest.mark.parametrize("name,steps,dt", [("channel_flux", 5, 1e-3), ("stretched", 4, 1e-3), ("custom", 4, 1e-3)])
def test_evolve_fp32_configurations(name, steps, dt):
    """The mixed form on the configurations whose (0,0) mode takes extra
    work: the constant-flux projection (ConstantMean, sources.jl:47-77) of the
    float64 (0,0) shadows, stretched levels, Neumann/Dirichlet boundary values
    in the float64 linear terms; SSPRK33 steps against the oracle."""
    ref, gpu = pair(name, precision="fp32", seed=12)
    O.evolve(ref, steps * dt, dt)
    B.evolve(gpu, steps * dt, dt)
    got = gpu.state
    for f in ref.fields:
        assert relerr(got[f], ref.state[f]) < 1e-4, (name, f, relerr(got[f], ref.state[f]))


@pytest.mark.parametrize("method", ["Euler", "SSPRK22", "AB2"])
def test_other_methods_fp32(method):
    ref, gpu = pair("les", precision="fp32", seed=9)
    O.evolve(ref, 4 * 5e-4, 5e-4, method=getattr(O, method)())
    B.evolve(gpu, 4 * 5e-4, 5e-4, method=getattr(B, method)())
    got = gpu.state
    for f in ref.fields:
        assert relerr(got[f], ref.state[f]) < 1e-4, (method, f, relerr(got[f], ref.state[f]))


@pytest.mark.parametrize("name", ["channel", "les_square"])
def test_ab2_history_survives_autotune(name):
    """The autotuner replays the first stage kernel, which writes utmp; the
    Adams–Bashforth method keeps f(u) of the previous step there, so autotune
    between AB2 steps must not change the next steps."""
    ref, gpu = pair(name, seed=13)
    dt = 1e-3 if name == "channel" else 5e-4
    O.evolve(ref, 6 * dt, dt, method=O.AB2())
    gpu.set_method(B.AB2())
    gpu.init_problem()
    t = gpu.step(dt, 3, t=0.0, t_end=6 * dt)
    gpu.autotune(reps=2)
    gpu.step(dt, 3, t=t, t_end=6 * dt)
    got = gpu.state
    for f in ref.fields:
        assert relerr(got[f], ref.state[f]) < RTOL64, (name, f, relerr(got[f], ref.state[f]))


def test_ab2_requires_constant_dt():
    ref, gpu = pair("channel")
    gpu.set_method(B.AB2())
    gpu.init_problem()
    gpu.step(1e-3, 2)
    with pytest.raises(B._lib.BldError, match="Cannot change time step"):
        gpu.step(2e-3, 1)


def test_constant_flux_keeps_the_mean():
    """On a uniform grid (weights summing to one) every projection leaves the
    z-weighted mean of the (0,0) mode of vel1 and vel2 at the prescribed flux
    (sources.jl:71-77).  The reference adds the same shift to every layer, so on
    a stretched grid, where the weights sum to one only approximately, a
    residual remains; the stretched case is covered by the parity tests."""
    mk = lambda M: M.Model((12, 10, 8), M.Domain((2 * np.pi, np.pi, 2.0), M.SmoothWall(), M.SmoothWall()),
                           M.incompressible_flow(1 / 180, constant_flux=(1.0, 0.2)))
    ref, gpu = mk(O), mk(B)
    gpu.set_state(**random_state(ref, np.random.default_rng(4)))
    B.evolve(gpu, 0.01, 1e-3)
    st = gpu.state
    w = [pd for k, pd in ref.disc if k == "mean"][0]["weight"]
    assert abs(np.sum(w) - 1.0) < 1e-14
    assert abs(np.sum(st["vel1"][0, 0, :].real * w) - 1.0) < 1e-12
    assert abs(np.sum(st["vel2"][0, 0, :].real * w) - 0.2) < 1e-12
