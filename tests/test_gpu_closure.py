"""Learned SGS closure on the GPU (through the C ABI) against its oracle, and
the filtered-DNS training pipeline: samples produced by the engine, training
on the GPU, and the a-priori comparison with Smagorinsky on held-out data.

Tolerances: the sampler runs in float64 and outputs float32 (2e-6 relative);
a constant-output network involves no bfloat16 rounding (float32 stress
arithmetic, 1e-5); a random network rounds its operands to bfloat16 on both
sides but accumulates in float32 on the GPU, so rare rounding flips of hidden
activations give ~1e-3 relative differences in the rates (5e-3 bound).
"""

import math

import numpy as np
import pytest

import oracle as O
import oracle.mlp as OM
from oracle import closure as OC
from conftest import load_package

pytestmark = pytest.mark.gpu

B = load_package()
CL = B.closure


def relerr(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def l2err(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def random_state(model, rng, amp=0.3, decay=0.4):
    K1, K2 = model.grid.fdsize()
    k1, k2 = O.wavenumbers(model.grid)
    scale = np.exp(-decay * np.sqrt(k1[:, None] ** 2 + k2[None, :] ** 2))[:, :, None]
    out = {}
    for f in model.fields:
        n = model.state[f].shape[2]
        a = (rng.standard_normal((K1, K2, n)) + 1j * rng.standard_normal((K1, K2, n))) * amp * scale
        km = model.grid.k2max
        a[0, K2 - km:, :] = np.conj(a[0, 1:km + 1, :][::-1, :])
        a[0, 0, :] = a[0, 0, :].real
        if f == "vel1":
            a[0, 0, :] += 2.0
        out[f] = np.asfortranarray(a)
    return out


def les_spec(M, n=(12, 10, 8), Cs=0.1, closure=None, guard=False):
    dom = M.Domain((2 * math.pi, 2 * math.pi, 1.0), M.RoughWall(1e-3, 0.4), M.FreeSlipBoundary())
    smag = M.StaticSmagorinskyModel(Cs=Cs, wall_damping=True)
    sgs = smag if closure is None else M.LearnedSGSModel(closure, fallback=smag, guard=guard)
    return n, dom, M.incompressible_flow(1e-3, constant_forcing=1, sgs_model=sgs)


def const_params(c):
    p = OM.init_params(0)
    for k in p:
        p[k] = np.zeros_like(p[k])
    p["b3"][:6] = c
    return p


def pair(params, precision="fp64", seed=3, guard=False, n=(12, 10, 8)):
    ref = O.Model(*les_spec(O, n))
    gpu = B.Model(*les_spec(B, n, closure=CL.pack(params), guard=guard), precision=precision)
    st = random_state(ref, np.random.default_rng(seed))
    gpu.set_state(**st)
    delta = math.sqrt((2 * math.pi / (2 * ref.grid.k1max + 2)) * (2 * math.pi / (2 * ref.grid.k2max + 2)))
    xw = OC.wall_feature(ref, delta)
    return ref, gpu, st, delta, xw


def test_sampler_matches_oracle():
    ref = O.Model(*les_spec(O, (16, 12, 8)))
    gpu = B.Model(*les_spec(B, (16, 12, 8)))
    st = random_state(ref, np.random.default_rng(5))
    gpu.set_state(**st)
    X, Y, delta = gpu.sample_closure(4)
    xw = OC.wall_feature(ref, delta)
    Xr, Yr = OC.sample(ref, st, 4, delta, xw)
    assert X.shape == Xr.shape == (4 * 3 * 8, 16)
    assert relerr(X, Xr) < 2e-6
    assert relerr(Y[:, :6], Yr[:, :6]) < 2e-6
    assert relerr(Y[:, 15], Yr[:, 15]) < 2e-6


def test_sampler_rejects_bad_filter():
    gpu = B.Model(*les_spec(B, (16, 12, 8)))
    with pytest.raises(B._lib.BldError):
        gpu.sample_closure(5)  # 16 is not a multiple of 5


@pytest.mark.parametrize("c", [[0.0] * 6, [0.3, -0.2, 0.05, 0.11, -0.07, 0.04]])
def test_constant_network_rates(c):
    params = const_params(np.array(c))
    ref, gpu, st, delta, xw = pair(params)
    want = OC.closure_rates(ref, st, delta, xw, lambda X: np.tile(np.array(c), (X.shape[0], 1)))
    got = gpu.rates()
    for f in ref.fields:
        assert relerr(got[f], want[f]) < 1e-5, (f, relerr(got[f], want[f]))


def test_zero_network_equals_smagorinsky_without_viscosity():
    ref, gpu, st, delta, xw = pair(const_params(np.zeros(6)))
    smag = B.Model(*les_spec(B, Cs=1e-14))
    smag.set_state(**st)
    a, b = gpu.rates(), smag.rates()
    for f in a:
        assert relerr(a[f], b[f]) < 1e-12


@pytest.mark.parametrize("guard", [False, True])
def test_random_network_rates(guard):
    params = OM.init_params(7)
    ref, gpu, st, delta, xw = pair(params, guard=guard)
    want = OC.closure_rates(ref, st, delta, xw, OC.network(params), guard=guard)
    got = gpu.rates()
    for f in ref.fields:
        assert l2err(got[f], want[f]) < 5e-3, (f, l2err(got[f], want[f]))


def test_closure_fp32_and_removal():
    params = OM.init_params(11)
    ref, g64, st, delta, xw = pair(params)
    _, g32, _, _, _ = pair(params, precision="fp32")
    a, b = g64.rates(), g32.rates()
    for f in a:
        assert l2err(b[f], a[f]) < 5e-3
    # params = None restores the Smagorinsky stresses
    smag = B.Model(*les_spec(B))
    smag.set_state(**st)
    g64._check(g64._lib.bld_sgs_closure(g64._h, None, 0.0, None, 0))
    a, b = g64.rates(), smag.rates()
    for f in a:
        assert relerr(a[f], b[f]) < 1e-12


def test_closure_requires_sgs_process():
    dom = B.Domain((1.0, 1.0, 1.0), B.SmoothWall(), B.SmoothWall())
    m = B.Model((8, 8, 4), dom, B.incompressible_flow(0.1))
    flat = CL.pack(OM.init_params(0))
    xw = np.zeros(4)
    with pytest.raises(B._lib.BldError):
        m._check(m._lib.bld_sgs_closure(m._h, B._lib.fptr(flat), 0.1, B._lib.dptr(xw), 0))


def _dns_samples(warmup=600, steps_between=100, snapshots=4, F=4, dt=5e-4):
    """Channel flow between smooth walls on a 64×48×32 grid (Re_τ = 180
    forcing): parabolic profile plus a smooth broadband perturbation, advanced
    on the GPU; box-filtered samples of several snapshots."""
    W = B.workloads
    n, dom, procs = W.dns_channel(B, (64, 48, 32))
    m = B.Model(n, dom, procs, precision="fp32")
    zc, zi = m.coordinates("vel1", 3), m.coordinates("vel3", 3)
    n1, n2 = m.default_size()
    u, v, w = W.dns_initial(zc, zi, n1, n2, seed=5, uc=10.0, amp=0.0)
    m.set_physical("vel1", u)
    m.set_physical("vel2", v)
    m.set_physical("vel3", w)
    st, noise = m.state, W.spectral_noise(m, seed=5, amp=0.25, k0=3.0)
    m.set_state(**{f: st[f] + noise[f] for f in st})
    m.init_problem()
    m.step(dt, warmup)
    xs, ts = [], []
    for _ in range(snapshots):
        m.step(dt, steps_between)
        x, t, _ = m.sample_closure(F)
        xs.append(x)
        ts.append(t)
    return xs, ts


def test_learned_closure_beats_smagorinsky_apriori():
    """Trained on filtered-DNS samples from the engine, the network predicts
    the held-out snapshot's SGS stress with a smaller error and a higher
    correlation than Smagorinsky, even with Smagorinsky's constant fitted to
    the held-out data."""
    xs, ts = _dns_samples()
    for t in ts:
        assert np.all(np.isfinite(t))
    xtr, ttr = np.concatenate(xs[:-1]), np.concatenate(ts[:-1])
    params, losses = CL.train(xtr, ttr, steps=1600, batch=4096, lr=3e-3, seed=1)
    assert losses[-20:].mean() < 0.5 * losses[:5].mean()
    res = CL.apriori(params, xs[-1], ts[-1])
    le, sb = res["learned"], res["smagorinsky_best"]
    assert le["rel_sq_error"] < 0.7 * sb["rel_sq_error"], res
    assert le["rel_sq_error"] < 1.0, res
    # the wall-normal shear stresses matter most for the mean flow
    for m in (4, 5):
        assert le["corr"][m] > sb["corr"][m] + 0.1, res


def test_learned_les_runs_and_stays_bounded():
    """A-posteriori: the LES with the trained closure (dissipation guard on)
    advances 200 steps from the same start as the Smagorinsky LES, stays
    finite, and its mean velocity profile stays within 10 % of Smagorinsky's
    (the two closures differ most in the first layers above the wall)."""
    xs, ts = _dns_samples(snapshots=2)
    params, _ = CL.train(np.concatenate(xs), np.concatenate(ts), steps=800, batch=4096, lr=3e-3, seed=2)
    W = B.workloads
    runs = {}
    for name in ("smag", "learned"):
        n, dom, procs = W.les_neutral(B, (32, 32, 16))
        if name == "learned":
            procs[-1] = B.LearnedSGSModel(CL.pack(params), fallback=procs[-1], guard=True)
        m = B.Model(n, dom, procs, precision="fp32")
        n1, n2 = m.default_size()
        u, v = W.les_initial(m.coordinates("vel1", 3), n1, n2, seed=3)
        m.set_physical("vel1", u)
        m.set_physical("vel2", v)
        m.set_physical("vel3", np.zeros((n1, n2, m.n3 - 1), order="F"))
        m.init_problem()
        m.step(5e-4 * 4, 200)
        st = m.state
        assert all(np.all(np.isfinite(a)) for a in st.values())
        runs[name] = st["vel1"][0, 0, :].real
    prof_s, prof_l = runs["smag"], runs["learned"]
    assert np.abs(prof_l - prof_s).max() < 0.10 * np.abs(prof_s).max()
