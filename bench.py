#!/usr/bin/env python3
"""Benchmark of the MI355X SSPRK33 time-step engine.

Default workload: the 128×128×64 neutral half-channel LES (rough wall z0=1e-4,
free slip top, static Smagorinsky with wall damping, SSPRK33, dt = 5e-4).  The
chosen form (float32 fields, float64 pressure elimination / (0,0) mode /
forcing) is timed against the plain float64 form on one GPU.  A step is one
SSPRK33 step (3 right-hand sides, 3 fused stage+projection passes), with the
state resident in HBM; steps are replayed from a captured HIP graph.

Prints ONE JSON line: steps/s of the chosen form, with the other legs as fields.
"""

import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(REPO, "boundarylayerdynamics.jl_amd")
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E peak, GB/s
METRIC = "steps/s & grid-pt updates/s, 128²×64 LES, chosen form vs FP64 plain, 1 GPU"


def load_package():
    if "bld_amd" in sys.modules:
        return sys.modules["bld_amd"]
    spec = importlib.util.spec_from_file_location(
        "bld_amd", os.path.join(PKG_DIR, "__init__.py"), submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["bld_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def make_les(B, precision, res):
    W = B.workloads
    n, dom, procs = W.les_neutral(B, res)
    m = B.Model(n, dom, procs, precision=precision)
    n1, n2 = m.default_size()
    u, v = W.les_initial(m.coordinates("vel1", 3), n1, n2)
    m.set_physical("vel1", u)
    m.set_physical("vel2", v)
    m.set_physical("vel3", np.zeros((n1, n2, m.n3 - 1), order="F"))
    m.init_problem(0.0)  # projection + first rate (ODEProblem)
    return m


def time_form(B, precision, res, dt, steps, warmup, profile_steps, autotune=False):
    m = make_les(B, precision, res)
    if autotune:
        m.autotune(reps=3)
    m.step(dt, warmup, graph=True)
    t0 = time.perf_counter()
    ms_events = m.time_steps(dt, steps, graph=True)
    wall = time.perf_counter() - t0
    kt = None
    if profile_steps:
        # an idle gap before the serialized profiled steps, so that a kernel
        # trace of this run can tell them from the graph replays (which run
        # the two streams concurrently)
        time.sleep(PROFILE_GAP_S)
        m.set_profiling(True)
        m.step(dt, profile_steps, graph=False)
        kt = m.kernel_times()
        m.set_profiling(False)
    st = m.state
    finite = all(np.all(np.isfinite(a)) for a in st.values())
    return m, dict(ms_per_step=1e3 * wall / steps, ms_per_step_events=ms_events / steps, finite=finite), kt


def roofline(B, m, kt, precision, res, adv_n, sgs_n):
    sizes = dict(K1=m.k1max + 1, K2=2 * m.k2max + 1, n3=m.n3,
                 spaces=[("adv", *adv_n), ("sgs", *sgs_n)])
    ab = B.workloads.algorithmic_bytes(sizes, None, precision, m.tuning()["transform_path"][0])
    # ab[k]: algorithmic bytes of kernel class k per right-hand side (all spaces);
    # a class may take several dispatches per right-hand side (one per space or
    # LDS class), so bytes and durations are averaged per dispatch
    n_rhs = kt["linear"][1]
    per = {}
    for k, (ms, n) in kt.items():
        if n == 0 or k not in ab:
            continue
        calls = n_rhs if k != "stage_project" else n
        total_bytes = ab[k] * calls
        per[k] = dict(launches=n, avg_us=1e3 * ms / n, total_ms=ms, bytes_per_launch=total_bytes / n,
                      gbs=total_bytes / (ms * 1e-3) / 1e9)
    dom = max(per, key=lambda k: per[k]["total_ms"])
    d = per[dom]
    avg_us, bpl, gbs = d["avg_us"], d["bytes_per_launch"], d["gbs"]
    # the dominant class's dispatches replayed back to back between two HIP
    # events (bld_time_kernel): the duration of a dispatch without the dispatch
    # latency that an event pair around every single dispatch adds
    try:
        us, nd = m.time_kernel(dom, reps=20)
    except B._lib.BldError:
        us, nd = 0.0, 0
    if us > 0 and nd > 0:
        avg_us, bpl = us, ab[dom] / nd
        gbs = bpl / (us * 1e-6) / 1e9
    return dict(kernel=dom, bound="hbm", achieved=round(gbs, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(gbs / HBM_PEAK_GBS, 4), traffic=None, avg_us=round(avg_us, 2),
                avg_us_event_pairs=round(d["avg_us"], 2), bytes_per_launch=int(bpl)), per


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo) and its logical core count."""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return f"{line.split(':', 1)[1].strip()} ({os.cpu_count()} logical cores visible)"
    except OSError:
        pass
    return None


def cpu_baseline(res, dt, budget_s):
    """The oracle (float64 NumPy/scipy.fft restatement of the reference) on a
    bounded sample of the same workload: a few SSPRK33 steps on the host cores."""
    sys.path.insert(0, REPO)
    import oracle as O
    import oracle.model as OM
    import scipy.fft  # noqa: F401

    B = load_package()
    W = B.workloads
    n, dom, procs = W.les_neutral(O, res)
    m = O.Model(n, dom, procs)
    n1, n2 = O.pdsize(m.grid)
    u, v = W.les_initial(m.coordinates("vel1", 3), n1, n2)
    for f in m.state.values():
        f[...] = 0
    O.add_field(m.state["vel1"], u)
    O.add_field(m.state["vel2"], v)
    cores = os.cpu_count() or 1
    OM.FFT_WORKERS = cores
    prob = O.ODEProblem(OM.rate_fn(m), [m.state[f] for f in m.fields], (0.0, 1e9),
                        projection=OM.projection_fn(m))
    buf = {"utmp": [np.zeros_like(x) for x in prob.u]}
    from oracle.ode import perform_step
    perform_step(prob, dt, O.SSPRK33(), buf)  # warm-up
    prob.rate(prob.du, prob.u, dt)
    nsteps, t0 = 0, time.perf_counter()
    while True:
        perform_step(prob, dt, O.SSPRK33(), buf)
        prob.rate(prob.du, prob.u, dt)
        nsteps += 1
        el = time.perf_counter() - t0
        if el > budget_s or nsteps >= 50:
            break
    return dict(value=round(nsteps / el, 4), unit="steps/s", cores=cores, cpu_model=cpu_model(), kind="port",
                sample=f"{nsteps} SSPRK33 steps (+1 warm-up) of the same {n[0]}x{n[1]}x{n[2]} LES in float64, "
                       f"oracle (NumPy + scipy.fft with {cores} workers), {el:.1f} s")


def make_dns(B, precision, res=(64, 64, 64)):
    W = B.workloads
    n, dom, procs = W.dns_channel(B, res)
    m = B.Model(n, dom, procs, precision=precision)
    n1, n2 = m.default_size()
    u, v, w = W.dns_initial(m.coordinates("vel1", 3), m.coordinates("vel3", 3), n1, n2)
    for f, vals in zip(("vel1", "vel2", "vel3"), (u, v, w)):
        m.set_physical(f, vals)
    m.init_problem(0.0)
    return m


def c1_bench(B, steps=200, warmup=10, profile_steps=5, cpu_seconds=8.0):
    """The 64³ closed-channel DNS in plain float64 (ν = 1/180, forcing 1,
    workflow parabola + 5 % seeded noise, SSPRK33 dt = 1e-3, 200 steps): steps/s,
    grid-point updates/s, the roofline of its dominant kernel class, and the
    oracle on a bounded sample of the same workload."""
    res, dt = (64, 64, 64), 1e-3
    m = make_dns(B, "fp64", res)
    m.autotune(reps=3)
    m.step(dt, warmup)
    ms = m.time_steps(dt, steps)
    time.sleep(PROFILE_GAP_S)
    # one unprofiled eager step first: the first launches after the idle gap
    # run at a lower clock and would decide which class is the dominant one
    m.step(dt, 1, graph=False)
    m.set_profiling(True)
    m.step(dt, profile_steps, graph=False)
    kt = m.kernel_times()
    m.set_profiling(False)
    adv_n = B.pdsize((m.k1max, m.k2max), "quadratic")
    sizes = dict(K1=m.k1max + 1, K2=2 * m.k2max + 1, n3=m.n3, spaces=[("adv", *adv_n)])
    ab = B.workloads.algorithmic_bytes(sizes, None, "fp64", m.tuning()["transform_path"][0])
    n_rhs = kt["linear"][1]
    per = {}
    for k, (kms, n) in kt.items():
        if n == 0 or k not in ab:
            continue
        calls = n_rhs if k != "stage_project" else n
        per[k] = dict(avg_us=1e3 * kms / n, total_ms=kms, bytes_per_launch=ab[k] * calls / n)
    dom = max(per, key=lambda k: per[k]["total_ms"])
    avg_us, bpl = per[dom]["avg_us"], per[dom]["bytes_per_launch"]
    try:
        us, nd = m.time_kernel(dom, reps=20)
        if us > 0 and nd > 0:
            avg_us, bpl = us, ab[dom] / nd
    except B._lib.BldError:
        pass
    gbs = bpl / (avg_us * 1e-6) / 1e9
    finite = all(np.all(np.isfinite(a)) for a in m.state.values())
    sps = 1e3 * steps / ms
    cpu = None
    if cpu_seconds > 0:
        sys.path.insert(0, REPO)
        import oracle as O
        import oracle.model as OM
        from oracle.ode import perform_step
        W = B.workloads
        n, dom_o, procs = W.dns_channel(O, res)
        r = O.Model(n, dom_o, procs)
        n1, n2 = O.pdsize(r.grid)
        u, v, w = W.dns_initial(r.coordinates("vel1", 3), r.coordinates("vel3", 3), n1, n2)
        for f, vals in zip(r.fields, (u, v, w)):
            r.state[f][...] = 0
            O.add_field(r.state[f], vals)
        OM.FFT_WORKERS = os.cpu_count() or 1
        prob = O.ODEProblem(OM.rate_fn(r), [r.state[f] for f in r.fields], (0.0, 1e9),
                            projection=OM.projection_fn(r))
        buf = {"utmp": [np.zeros_like(x) for x in prob.u]}
        perform_step(prob, dt, O.SSPRK33(), buf)
        prob.rate(prob.du, prob.u, dt)
        k, t0 = 0, time.perf_counter()
        while True:
            perform_step(prob, dt, O.SSPRK33(), buf)
            prob.rate(prob.du, prob.u, dt)
            k += 1
            el = time.perf_counter() - t0
            if el > cpu_seconds or k >= 200:
                break
        cpu = dict(value=round(k / el, 3), unit="steps/s", cores=os.cpu_count(), kind="port",
                   sample=f"{k} SSPRK33 steps (+1 warm-up) of the 64x64x64 DNS in float64, oracle, {el:.1f} s")
    return dict(workload="closed-channel DNS 64x64x64 fp64, SSPRK33 dt=1e-3", steps=steps,
                steps_per_s=round(sps, 2), ms_per_step=round(ms / steps, 4),
                grid_point_updates_per_s=round(sps * res[0] * res[1] * res[2], 1), finite=finite,
                transform_path=m.tuning()["transform_path"],
                roofline=dict(kernel=dom, bound="hbm", achieved=round(gbs, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                              frac=round(gbs / HBM_PEAK_GBS, 4), avg_us=round(avg_us, 2), bytes_per_launch=int(bpl)),
                step_bytes_fraction=round(17 * 16 * sizes["K1"] * sizes["K2"] * (3 * m.n3 - 1) * sps / 8e12, 4),
                kernels={k: round(v["avg_us"], 2) for k, v in per.items()}, cpu_baseline=cpu)


def _arr_of(disc, i, n):
    """The i-th float64 array the configuration of `disc` points to (0: α at
    C-nodes [n3], 1: α at I-nodes incl. the walls [n3+1])."""
    a = disc._keep[i]
    assert a.shape == (n,)
    return a


def pressure_of(B, m, state, rates, dt):
    """The pressure that the first stage of the next SSPRK33 step solves for:
    the projection of u + dt·f(u) (pressure.jl:300-426), from the state and
    rates of `m` read back in float64, solved with the engine's batched LDLt
    in float64.  The scaled pressure of the reference (αc⁻¹-scaled system)."""
    d = m.disc
    n3, K1, K2 = m.n3, m.k1max + 1, 2 * m.k2max + 1
    ac, ai = _arr_of(d, 0, n3), _arr_of(d, 1, n3 + 1)
    L1, L2 = m.domain.hsize
    k1 = np.arange(K1)
    k2 = np.concatenate([np.arange(m.k2max + 1), np.arange(-m.k2max, 0)])
    D1 = (1j * 2 * np.pi * k1 / L1)[:, None, None]
    D2 = (1j * 2 * np.pi * k2 / L2)[None, :, None]
    u, v, w = (state[f] + dt * rates[f] for f in ("vel1", "vel2", "vel3"))
    z = np.zeros((K1, K2, 1), complex)
    rhs = (D1 * u + D2 * v) / ac[None, None, :] + np.concatenate([w, z], axis=2) - np.concatenate([z, w], axis=2)
    rhs[0, 0, -1] = 0  # walls at w = 0; the (0,0) row pinned
    d3g3 = np.zeros(n3)
    d3g3[1:] -= ai[1:n3]
    d3g3[:-1] -= ai[1:n3]
    kk1, kk2 = np.tile(k1, K2), np.repeat(k2, K1)
    shift = 4 * np.pi ** 2 * (kk1 ** 2 / L1 ** 2 + kk2 ** 2 / L2 ** 2)
    dvs = d3g3[None, :] - shift[:, None] / ac[None, :]
    dvs[0] = d3g3
    dvs[0, -1] = -3 * ai[n3 - 1]
    evs = np.broadcast_to(ai[1:n3], (K1 * K2, n3 - 1)).copy()
    b = rhs.reshape((K1 * K2, n3), order="F")
    return B.BatchLDLt(dvs, evs, precision="fp64").solve(b)


def form_error(B, res, dt, checkpoints=(100, 500)):
    """Error of the chosen (mixed) form against the plain float64 form on the
    GPU, both from the same projected state (log law + seeded noise): per
    field the relative L2 and max-norm difference after each checkpoint's
    number of SSPRK33 steps, plus that of the pressure the next stage solves
    for.  Relative to the float64 field's L2 norm / maximum."""
    m64 = make_les(B, "fp64", res)
    st0 = m64.state
    m32 = B.Model(*B.workloads.les_neutral(B, res), precision="fp32")
    m32.set_state(**st0)
    m32.init_problem(0.0)
    out, done = {}, 0
    for n in checkpoints:
        m64.step(dt, n - done)
        m32.step(dt, n - done)
        done = n
        a, b = m32.state, m64.state
        e = {}
        for f in ("vel1", "vel2", "vel3"):
            d = a[f] - b[f]
            e[f] = dict(rel_l2=float(np.linalg.norm(d) / np.linalg.norm(b[f])),
                        rel_max=float(np.abs(d).max() / np.abs(b[f]).max()))
        p32 = pressure_of(B, m32, a, m32.rates(), dt)
        p64 = pressure_of(B, m64, b, m64.rates(), dt)
        d = p32 - p64
        e["pressure"] = dict(rel_l2=float(np.linalg.norm(d) / np.linalg.norm(p64)),
                             rel_max=float(np.abs(d).max() / np.abs(p64).max()))
        out[f"steps_{n}"] = {f: {k: float(f"{x:.3e}") for k, x in v.items()} for f, v in e.items()}
    return out


def pressure_solve(B, nz=256, n=256, reps=5):
    """Isolated batched pressure solve (65,536 modes × nz), float64 and
    float32 + one refinement step; bytes per solve = 32 B per unknown."""
    K1, K2 = n, n
    k1 = np.arange(K1)
    k2 = np.concatenate([np.arange(K2 // 2), np.arange(-(K2 // 2), 0)])
    a = float(nz)
    d3g3 = np.full(nz, -2 * a)
    d3g3[0] = d3g3[-1] = -a
    kk1, kk2 = np.tile(k1, K2), np.repeat(k2, K1)
    shift = 4 * np.pi ** 2 * (kk1 ** 2 + kk2 ** 2) / (2 * np.pi) ** 2
    dvs = d3g3[None, :] - shift[:, None] / a
    dvs[0] = d3g3
    dvs[0, -1] = -3 * a
    evs = np.full((K1 * K2, nz - 1), a)
    rng = np.random.Generator(np.random.PCG64(3))
    b = rng.standard_normal(dvs.shape) + 1j * rng.standard_normal(dvs.shape)
    b[0, -1] = 0
    out = {}
    x64 = None
    for prec in ("fp64", "fp32"):
        s = B.BatchLDLt(dvs, evs, precision=prec)
        ms, x = s.bench(b, reps=reps)
        if prec == "fp64":
            x64 = x
            err = None
        else:
            err = float(np.abs(x - x64).max() / np.abs(x64).max())
        byts = 32.0 * dvs.size
        out[prec] = dict(ms_per_solve=round(ms, 4), gbs=round(byts / (ms * 1e-3) / 1e9, 1),
                         rel_err_vs_fp64=err)
    out["systems"] = int(K1 * K2)
    out["nz"] = nz
    return out


def ensemble_bench(B, dt, steps, warmup, autotune):
    """32 independent LES models of uneven grids (8 each of 48²×32, 64²×48,
    96²×64, 128²×96) with seeded noise and Cs in [0.08, 0.16], fp32 fields,
    advanced together; one HIP graph per ensemble step."""
    W = B.workloads
    rng = np.random.Generator(np.random.PCG64(17))
    specs, grids = [], [(48, 48, 32), (64, 64, 48), (96, 96, 64), (128, 128, 96)]
    for g in grids:
        for _ in range(8):
            specs.append(W.les_neutral(B, g, Cs=float(rng.uniform(0.08, 0.16)), fft_friendly=True))
    ens = B.Ensemble(specs, precision="fp32")
    for i, m in enumerate(ens.members):
        n1, n2 = m.default_size()
        u, v = W.les_initial(m.coordinates("vel1", 3), n1, n2, seed=1000 + i)
        m.set_physical("vel1", u)
        m.set_physical("vel2", v)
        m.set_physical("vel3", np.zeros((n1, n2, m.n3 - 1), order="F"))
    ens.init_problem()
    if autotune:
        ens.autotune(reps=3)
    ens.step(dt, warmup)
    ms = ens.time_steps(dt, steps)
    ens.set_profiling(True)
    ens.time_steps(dt, 5, graph=False)
    kt = {k: round(v[0] / 5, 4) for k, v in ens.kernel_times().items() if v[1]}
    ens.set_profiling(False)
    finite = all(np.all(np.isfinite(a)) for m in ens.members[::8] for a in m.state.values())
    sps = 1e3 * steps / ms
    pts = sum(g[0] * g[1] * g[2] * 8 for g in grids)
    adv = sorted({tuple(s[2][0].dealiasing) for s in specs})
    return dict(instances=len(ens), grid_points=pts, advection_grids=[list(a) for a in adv], ensemble_steps_per_s=round(sps, 2),
                ms_per_ensemble_step=round(ms / steps, 4), grid_point_updates_per_s=round(sps * pts, 1),
                autotuned=autotune, finite=finite, kernel_ms_per_step=kt)


def ensemble_identical(B, dt, steps, warmup, copies=4, res=(128, 128, 64)):
    """`copies` independent 128x128x64 LES (the headline model, seeded noise
    per member) advanced as one ensemble: one launch per kernel class covers
    all members.  Grid-point updates per second of the ensemble, to set
    against the single LES."""
    W = B.workloads
    specs = [W.les_neutral(B, res) for _ in range(copies)]
    ens = B.Ensemble(specs, precision="fp32")
    for i, m in enumerate(ens.members):
        n1, n2 = m.default_size()
        u, v = W.les_initial(m.coordinates("vel1", 3), n1, n2, seed=2000 + i)
        m.set_physical("vel1", u)
        m.set_physical("vel2", v)
        m.set_physical("vel3", np.zeros((n1, n2, m.n3 - 1), order="F"))
    ens.init_problem()
    ens.autotune(reps=3)
    ens.step(dt, warmup)
    ms = ens.time_steps(dt, steps)
    sps = 1e3 * steps / ms
    finite = all(np.all(np.isfinite(a)) for m in ens.members for a in m.state.values())
    return dict(instances=copies, grid=list(res), ensemble_steps_per_s=round(sps, 2),
                ms_per_ensemble_step=round(ms / steps, 4),
                grid_point_updates_per_s=round(sps * copies * res[0] * res[1] * res[2], 1), finite=finite)


MLP_FLOP_PER_POINT = 2 * (10 * 64 + 64 * 64 + 64 * 6)  # forward, useful (unpadded) multiply-adds
BF16_DENSE_PEAK_TFLOPS = 2500.0


TRAINED_CLOSURE = os.path.join(PKG_DIR, "closure_params_dns256.npy")


def closure_params(B):
    """The network trained on the 256x256x128 channel DNS (train_closure.py),
    or an untrained one if that file is absent: (flat float32 params, label)."""
    if os.path.exists(TRAINED_CLOSURE):
        return np.load(TRAINED_CLOSURE).astype(np.float32), "trained on the 256x256x128 fp64 channel DNS"
    return B.closure.pack(B.closure.init_params(3)), "untrained (seeded initial weights)"


PROFILE_NAMES = ("vel1", "vel2", "sgs11", "sgs22", "sgs33", "sgs12", "sgs13", "sgs23")


def closure_vs_smagorinsky(B, params, res, dt, steps):
    """The LES with the learned closure compared with the same LES with
    Smagorinsky: both from the same seeded state, `steps` SSPRK33 steps with
    the mean profiles of the reference's MeanProfiles collected on the device;
    the relative L2 difference of the time-mean profiles (velocity and the
    SGS stresses) and the mean rough-wall stress of each run."""
    W = B.workloads
    means = {}
    for kind in ("smagorinsky", "learned"):
        n, dom, procs = W.les_neutral(B, res)
        if kind == "learned":
            procs[-1] = B.LearnedSGSModel(params, fallback=procs[-1], guard=True)
        m = B.Model(n, dom, procs, precision="fp32")
        n1, n2 = m.default_size()
        u, v = W.les_initial(m.coordinates("vel1", 3), n1, n2)
        m.set_physical("vel1", u)
        m.set_physical("vel2", v)
        m.set_physical("vel3", np.zeros((n1, n2, m.n3 - 1), order="F"))
        m.init_problem()
        m.enable_profiles(PROFILE_NAMES)
        m.step(dt, steps)
        integ, k = m.read_profiles()
        means[kind] = {p: integ[p] / (k * dt) for p in PROFILE_NAMES}
        means[kind]["finite"] = all(np.all(np.isfinite(a)) for a in m.state.values())
        del m
    s, l = means["smagorinsky"], means["learned"]
    rel = lambda a, b: float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
    return dict(steps=steps, time_span=steps * dt, finite=bool(s["finite"] and l["finite"]),
                mean_profile_rel_l2={p: round(rel(l[p], s[p]), 5) for p in PROFILE_NAMES},
                wall_stress_sgs13={"smagorinsky": float(s["sgs13"][0]), "learned": float(l["sgs13"][0])},
                resolved_vs_sgs13_top_half={"smagorinsky": float(np.mean(s["sgs13"][len(s["sgs13"]) // 2:])),
                                            "learned": float(np.mean(l["sgs13"][len(l["sgs13"]) // 2:]))})


def closure_bench(B, res, dt, steps, warmup, versus=True):
    """The same LES with the learned SGS closure (bf16 MFMA network at every
    C-node, Smagorinsky fallback and dissipation guard) in fp32, plus the
    network's standalone forward / training-gradient throughput and (with
    `versus`) the comparison with the Smagorinsky LES."""
    W, CL = B.workloads, B.closure
    params, label = closure_params(B)
    n, dom, procs = W.les_neutral(B, res)
    procs[-1] = B.LearnedSGSModel(params, fallback=procs[-1], guard=True)
    m = B.Model(n, dom, procs, precision="fp32")
    n1, n2 = m.default_size()
    u, v = W.les_initial(m.coordinates("vel1", 3), n1, n2)
    m.set_physical("vel1", u)
    m.set_physical("vel2", v)
    m.set_physical("vel3", np.zeros((n1, n2, m.n3 - 1), order="F"))
    m.init_problem()
    m.autotune(reps=3)
    m.step(dt, warmup)
    ms = m.time_steps(dt, steps)
    m.set_profiling(True)
    m.time_steps(dt, 5, graph=False)
    kt = m.kernel_times()
    m.set_profiling(False)
    finite = all(np.all(np.isfinite(a)) for a in m.state.values())
    ms_mlp, launches = kt["sgs_mlp"]
    pts = n1 * n2 * m.n3
    tflops = MLP_FLOP_PER_POINT * pts / (ms_mlp / launches * 1e-3) / 1e12
    net = CL.MLP(seed=1)
    npts = 1 << 20
    f_ms, g_ms = net.time(npts, reps=10)
    path = m.tuning()["transform_path"]
    del m
    versus = closure_vs_smagorinsky(B, params, res, dt, 400) if versus else None
    return dict(network=label, steps_per_s=round(1e3 * steps / ms, 2), ms_per_step=round(ms / steps, 4), finite=finite,
                transform_path=path,
                sgs_mlp_avg_us=round(1e3 * ms_mlp / launches, 2), points_per_launch=pts,
                sgs_mlp_tflops=round(tflops, 2), sgs_mlp_frac_bf16_peak=round(tflops / BF16_DENSE_PEAK_TFLOPS, 4),
                kernels={k: round(1e3 * v[0] / v[1], 2) for k, v in kt.items() if v[1]},
                mlp_forward_1M_ms=round(f_ms, 4),
                mlp_forward_tflops=round(MLP_FLOP_PER_POINT * npts / (f_ms * 1e-3) / 1e12, 2),
                mlp_gradient_1M_ms=round(g_ms, 4),
                mlp_gradient_tflops=round(3 * MLP_FLOP_PER_POINT * npts / (g_ms * 1e-3) / 1e12, 2),
                versus_smagorinsky=versus)


def copy_probe(mib=256, reps=10):
    """A plain device-to-device copy of known size (torch's copy kernel), the
    scale for the FETCH_SIZE / WRITE_SIZE counters: each dispatch reads and
    writes exactly `mib` MiB."""
    import torch
    n = mib * 1024 * 1024 // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": n * 4, "reps": reps}), flush=True)


def _csv_rows(directory, suffix):
    import csv
    import glob
    rows = []
    for path in sorted(glob.glob(os.path.join(directory, "**", "*" + suffix), recursive=True)):
        with open(path, newline="") as fh:
            rows.extend(csv.DictReader(fh))
    return rows


def _short(name):
    return name.replace("void ", "").replace("bld::", "").split("(")[0]


def _last_half(rows, key):
    """The later half of a kernel's dispatches (the timed steps come after the
    autotuner's trial launches and make up most of the run)."""
    rows = sorted(rows, key=key)
    return rows[len(rows) // 2:]


def _counter_per_kernel(directory, counter):
    """Mean of `counter` (KiB per dispatch) per kernel over the later half of
    its dispatches."""
    rows = [r for r in _csv_rows(directory, "counter_collection.csv") if r.get("Counter_Name") == counter]
    by = {}
    for r in rows:
        by.setdefault(_short(r["Kernel_Name"]), []).append(r)
    out = {}
    for k, rs in by.items():
        vals = [float(r["Counter_Value"]) for r in _last_half(rs, lambda r: int(r.get("Dispatch_Id", 0) or 0))]
        out[k] = (sum(vals) / len(vals), len(vals))
    return out


PROFILE_GAP_S = 0.05


def _after_last_gap(trace, gap_ns):
    """Dispatches after the last idle gap of at least gap_ns in the trace."""
    ev = sorted(trace, key=lambda r: int(r["Start_Timestamp"]))
    end, cut = None, 0
    for i, r in enumerate(ev):
        s0 = int(r["Start_Timestamp"])
        if end is not None and s0 - end >= gap_ns:
            cut = i
        end = max(end or 0, int(r["End_Timestamp"]))
    return ev[cut:]


def rocprof_summary(kt_dir, fetch_dir, write_dir, copy_fetch_dir, copy_write_dir, copy_bytes):
    """Markdown summary of a rocprofv3 --kernel-trace --stats run and of the
    FETCH_SIZE / WRITE_SIZE runs, the counters scaled by what they read for the
    copy probe (copy_bytes read and written per dispatch)."""
    lines = ["| kernel | calls | average us | total us | % |", "|---|---|---|---|---|"]
    for r in sorted(_csv_rows(kt_dir, "kernel_stats.csv"), key=lambda r: -float(r["TotalDurationNs"])):
        lines.append(f"| {_short(r['Name'])} | {r['Calls']} | {float(r['AverageNs']) / 1e3:.2f} | "
                     f"{float(r['TotalDurationNs']) / 1e3:.1f} | {float(r['Percentage']):.1f} |")
    # per-dispatch trace: the timed configuration only (grid size of each
    # kernel's last dispatch), without the autotuner's trial launches
    trace = _csv_rows(kt_dir, "kernel_trace.csv")
    by = {}
    for r in trace:
        by.setdefault(_short(r["Kernel_Name"]), []).append(r)
    lines += ["", "Timed configuration (the later half of each kernel's dispatches, after the autotuner's "
                  "trial launches):", "",
              "| kernel | dispatches | grid sizes | average us |", "|---|---|---|---|"]
    grid_of = lambda r: r.get("Grid_Size") or r.get("Grid_Size_X")
    for k, rs in sorted(by.items(), key=lambda kv: -sum(int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
                                                         for r in kv[1])):
        late = _last_half(rs, lambda r: int(r["Start_Timestamp"]))
        d = [int(r["End_Timestamp"]) - int(r["Start_Timestamp"]) for r in late]
        grids = sorted({int(grid_of(r)) for r in late})
        lines.append(f"| {k} | {len(d)} | {' '.join(map(str, grids))} | {sum(d) / len(d) / 1e3:.2f} |")
    # the serialized profiled steps that follow bench.py's idle gap: the
    # per-dispatch durations behind bench.py's kernel figures and roofline
    prof = _after_last_gap(trace, int(0.8 * PROFILE_GAP_S * 1e9))
    byp = {}
    for r in prof:
        byp.setdefault(_short(r["Kernel_Name"]), []).append(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
    lines += ["", "Profiled steps (serialized on one stream, after bench.py's idle gap: the profiled steps "
                  "behind bench.py's per-kernel averages, then its back-to-back replay of the dominant class "
                  "for the roofline):", "",
              "| kernel | dispatches | average us |", "|---|---|---|"]
    for k, d in sorted(byp.items(), key=lambda kv: -sum(kv[1])):
        lines.append(f"| {k} | {len(d)} | {sum(d) / len(d) / 1e3:.2f} |")
    cf = _counter_per_kernel(copy_fetch_dir, "FETCH_SIZE")
    cw = _counter_per_kernel(copy_write_dir, "WRITE_SIZE")
    ck = max(cf, key=lambda k: cf[k][0])
    kib = copy_bytes / 1024
    sf, sw = kib / cf[ck][0], kib / cw[max(cw, key=lambda k: cw[k][0])][0]
    lines += ["", f"Copy probe ({copy_bytes / 2 ** 20:.0f} MiB read + written per dispatch, kernel `{ck}`): "
                  f"FETCH_SIZE {cf[ck][0]:.0f} KiB, scale {sf:.4f}; WRITE_SIZE "
                  f"{cw[max(cw, key=lambda k: cw[k][0])][0]:.0f} KiB, scale {sw:.4f}.", "",
              "| kernel | dispatches | FETCH_SIZE KiB | WRITE_SIZE KiB | traffic MB (scaled) |", "|---|---|---|---|---|"]
    f = _counter_per_kernel(fetch_dir, "FETCH_SIZE")
    w = _counter_per_kernel(write_dir, "WRITE_SIZE")
    for k in sorted(f, key=lambda k: -f[k][0]):
        if k in w:
            mb = (f[k][0] * sf + w[k][0] * sw) * 1024 / 1e6
            lines.append(f"| {k} | {f[k][1]} | {f[k][0]:.0f} | {w[k][0]:.0f} | {mb:.2f} |")
    print("\n".join(lines))


TRAFFIC_FILE = os.path.join(REPO, "profiles", "traffic_les128x128x64_fp32.md")


KERNELS_FILE = os.path.join(REPO, "profiles", "les128x128x64_fp32_kernels.md")


def _md_rows(path, ncells):
    """The table rows with ncells cells of a markdown file."""
    out = []
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if line.startswith("|") and len(cells) == ncells:
            out.append(cells)
    return out


def measured_traffic(kernel):
    """HBM bytes per dispatch of kernel class `kernel` (float32 kernels) from the
    committed counter summary (the 'traffic MB (scaled)' column): the mean over
    the class's kernels (a class such as c2r_2d has one kernel per block size)
    weighted by their dispatches in the profiled steps of the committed kernel
    summary, so that kernels only the autotuner tried do not count; or None."""
    if not (os.path.exists(TRAFFIC_FILE) and os.path.exists(KERNELS_FILE)):
        return None
    names = {"c2r_2d": ("k_c2r_2d<float", "k_c2r_2h<float")}.get(kernel, ("k_" + kernel + "<float",))
    text = open(KERNELS_FILE).read()
    if "Profiled steps" not in text:
        return None
    prof = text[text.index("Profiled steps"):]
    weight = {}
    for line in prof.splitlines():
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if line.startswith("|") and len(cells) == 3 and cells[0].startswith(names):
            try:
                weight[cells[0]] = int(cells[1])
            except ValueError:
                pass
    tot, n = 0.0, 0
    for cells in _md_rows(TRAFFIC_FILE, 5):
        if cells[0] in weight:
            try:
                tot += float(cells[4]) * 1e6 * weight[cells[0]]
                n += weight[cells[0]]
            except ValueError:
                return None
    return int(tot / n) if n else None


def counter_summary(directory):
    """Markdown table of every counter of a rocprofv3 --pmc run: the mean per
    dispatch of each kernel over the later half of its dispatches."""
    names = sorted({r["Counter_Name"] for r in _csv_rows(directory, "counter_collection.csv")})
    per = {n: _counter_per_kernel(directory, n) for n in names}
    kernels = sorted({k for v in per.values() for k in v})
    lines = ["| kernel | dispatches | " + " | ".join(names) + " |", "|---|---|" + "---|" * len(names)]
    for k in kernels:
        cnt = max(per[n][k][1] for n in names if k in per[n])
        vals = [f"{per[n][k][0]:.6g}" if k in per[n] else "" for n in names]
        lines.append(f"| {k} | {cnt} | " + " | ".join(vals) + " |")
    print("\n".join(lines))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--dt", type=float, default=5e-4)
    ap.add_argument("--res", type=str, default="128,128,64")
    ap.add_argument("--profile-steps", type=int, default=10)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-pressure", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-closure", action="store_true")
    ap.add_argument("--no-c1", action="store_true")
    ap.add_argument("--autotune", type=int, default=1)
    ap.add_argument("--copy-probe", action="store_true", help="only run the counter-scale copy kernel")
    ap.add_argument("--pressure-only", action="store_true",
                    help="only the isolated pressure solve (65,536 systems x 256 levels), for profiling")
    ap.add_argument("--closure-only", action="store_true",
                    help="only the learned-closure LES leg without the Smagorinsky comparison, for profiling")
    ap.add_argument("--counter-summary", metavar="DIR", help="summarise the counters of a rocprofv3 --pmc run")
    ap.add_argument("--rocprof-summary", nargs=5, metavar="DIR",
                    help="summarise rocprofv3 output: kernel-trace, FETCH_SIZE, WRITE_SIZE, copy FETCH, copy WRITE")
    args = ap.parse_args()
    if args.rocprof_summary:
        rocprof_summary(*args.rocprof_summary, copy_bytes=256 * 2 ** 20)
        return
    if args.counter_summary:
        counter_summary(args.counter_summary)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the engine has no multi-GPU path "
              f"in this library", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        copy_probe()
        return
    if args.pressure_only:
        print(json.dumps({"pressure_solve_65536x256": pressure_solve(load_package(), reps=20)}), flush=True)
        return
    res = tuple(int(x) for x in args.res.split(","))
    B = load_package()
    if args.closure_only:
        print(json.dumps({"learned_closure_les": closure_bench(B, res, args.dt, max(args.steps // 3, 10), 5,
                                                               versus=False)}), flush=True)
        return
    log("fp32 mixed form ...")
    m32, r32, kt32 = time_form(B, "fp32", res, args.dt, args.steps, args.warmup, args.profile_steps,
                               autotune=bool(args.autotune))
    tuning32 = m32.tuning() if args.autotune else None
    adv_n = B.pdsize((m32.k1max, m32.k2max), "quadratic")
    sgs_n = B.pdsize((m32.k1max, m32.k2max), None)
    roof, per32 = roofline(B, m32, kt32, "fp32", res, adv_n, sgs_n)
    if res == (128, 128, 64):
        roof["traffic"] = measured_traffic(roof["kernel"])
    log("fp32:", r32, "dominant", roof["kernel"])
    fp64 = None
    if not args.no_fp64:
        log("fp64 plain form ...")
        m64, r64, kt64 = time_form(B, "fp64", res, args.dt, max(args.steps // 3, 10), max(args.warmup // 3, 3),
                                   args.profile_steps, autotune=bool(args.autotune))
        roof64, per64 = roofline(B, m64, kt64, "fp64", res, adv_n, sgs_n)
        fp64 = dict(steps_per_s=round(1e3 / r64["ms_per_step"], 2), ms_per_step=round(r64["ms_per_step"], 4),
                    finite=r64["finite"], dominant_kernel=roof64["kernel"],
                    kernels={k: dict(avg_us=round(v["avg_us"], 2), gbs=round(v["gbs"], 1)) for k, v in per64.items()})
        del m64
        log("fp64:", r64)
    c1 = None
    if not args.no_c1:
        log("C1 64^3 DNS fp64 ...")
        c1 = c1_bench(B, cpu_seconds=0.0 if args.no_cpu else 8.0)
        log("c1:", c1)
    ferr = None
    if not args.no_fp64:
        log("fp32 vs fp64 over 500 steps ...")
        ferr = form_error(B, res, args.dt)
        log("form error:", ferr)
    pres = None if args.no_pressure else pressure_solve(B)
    ens = None
    if not args.no_ensemble:
        log("ensemble ...")
        ens = ensemble_bench(B, args.dt, max(args.steps // 10, 10), 5, bool(args.autotune))
        log("ensemble:", ens)
    ens4 = None
    if not args.no_ensemble and res == (128, 128, 64):
        log("ensemble of 4 headline LES ...")
        ens4 = ensemble_identical(B, args.dt, max(args.steps // 5, 10), 5)
        ens4["batching_gain_vs_single"] = round(ens4["grid_point_updates_per_s"] /
                                                (1e3 / r32["ms_per_step"] * res[0] * res[1] * res[2]), 3)
        log("ensemble x4:", ens4)
    clo = None
    if not args.no_closure:
        log("learned closure ...")
        clo = closure_bench(B, res, args.dt, max(args.steps // 3, 10), 5)
        log("closure:", clo)
    cpu = None if args.no_cpu else cpu_baseline(res, args.dt, args.cpu_seconds)
    sps = 1e3 / r32["ms_per_step"]
    points = res[0] * res[1] * res[2]
    line = {
        "metric": METRIC,
        "value": round(sps, 2),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r32["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (f64 pressure elimination, (0,0) mode, forcing)",
        "data": "synthetic: log-law profile + seeded Gaussian noise (sigma = 0.1 u*), projected",
        "config": {"workload": f"neutral half-channel LES {res[0]}x{res[1]}x{res[2]}, rough wall z0=1e-4, "
                               f"free-slip top, static Smagorinsky Cs=0.1 with wall damping, SSPRK33 dt={args.dt}",
                   "form": "fp32 fields + fp64 pressure/mean mode (chosen form)",
                   "advection_grid": list(adv_n), "sgs_grid": list(sgs_n), "hip_graph": True,
                   "fft_autotuned": bool(args.autotune),
                   "parallelism": "single GPU"},
        "grid_point_updates_per_s": round(sps * points, 1),
        "ms_per_step_hip_events": round(r32["ms_per_step_events"], 4),
        "finite": r32["finite"],
        "fp64_plain": fp64,
        "speedup_vs_fp64_plain": round(sps / fp64["steps_per_s"], 3) if fp64 else None,
        "kernels_fp32": {k: dict(avg_us=round(v["avg_us"], 2), launches=v["launches"], gbs=round(v["gbs"], 1))
                         for k, v in per32.items()},
        "fp32_vs_fp64_500": ferr,
        "dns_64cubed_fp64": c1,
        "roofline": roof,
        "fft_autotune": tuning32,
        "pressure_solve_65536x256": pres,
        "ensemble_32_uneven_les": ens,
        "ensemble_4_headline_les": ens4,
        "learned_closure_les": clo,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
