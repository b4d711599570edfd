"""Benchmark and parity workloads (the reference's channel/LES configurations).

Each workload is a domain + process list + seeded initial condition, built
against a module that provides the reference's model API (this package, or
any object with the same names), so the same specification can be run by the
GPU engine and by a CPU implementation.
"""

import math

import numpy as np


def les_neutral(M, n=(128, 128, 64), Cs=0.1, fft_friendly=False):
    """Neutral atmospheric half channel LES: 2π×2π×1, rough wall z0=1e-4, κ=0.4,
    free-slip top, static Smagorinsky with Mason wall damping, ν=1e-5,
    constant mean pressure gradient forcing 1 (les_test.jl:55-58 style).

    ``fft_friendly`` sets the advection grid through the reference's tuple
    dealiasing (physical_space.jl:45) to the smallest 2^a·3^b ≥ 3·kmax+1
    (alias-free for the quadratic products) instead of the `:quadratic`
    heuristic, which gives sizes such as 142 = 2·71 for kmax = 47."""
    dom = M.Domain((2 * math.pi, 2 * math.pi, 1.0), M.RoughWall(1e-4, 0.4), M.FreeSlipBoundary())
    procs = M.incompressible_flow(1e-5, constant_forcing=1,
                                  sgs_model=M.StaticSmagorinskyModel(Cs=Cs, wall_damping=True))
    if fft_friendly:
        sizes = tuple(smooth_size(3 * ((k - 1) // 2) + 1) for k in n[:2])
        procs[0] = M.MomentumAdvection(dealiasing=sizes)
    return n, dom, procs


def smooth_size(nmin):
    """Smallest 2^a·3^b ≥ nmin."""
    best = None
    a = 1
    while a < 4 * nmin:
        b = a
        while b < 4 * nmin:
            if b >= nmin and (best is None or b < best):
                best = b
            b *= 3
        a *= 2
    return best


def les_initial(domain_z, n1, n2, seed=7, z0=1e-4, kappa=0.4, sigma=0.1):
    """Log-law profile u = ln(z/z0)/κ (u* = 1) + seeded Gaussian noise σ on
    vel1 and vel2; vel3 = 0.  Arrays at the default physical size."""
    zc = np.asarray(domain_z)
    rng = np.random.Generator(np.random.PCG64(seed))
    shape = (n1, n2, len(zc))
    u = np.log(zc / z0)[None, None, :] / kappa + sigma * rng.standard_normal(shape)
    v = sigma * rng.standard_normal(shape)
    return np.asfortranarray(u), np.asfortranarray(v)


def dns_channel(M, n=(64, 64, 64)):
    """Closed channel DNS: 2π×π×2, smooth walls, ν = 1/180, constant forcing 1."""
    dom = M.Domain((2 * math.pi, math.pi, 2.0), M.SmoothWall(), M.SmoothWall())
    return n, dom, M.incompressible_flow(1 / 180, constant_forcing=1)


def dns_initial(zc, zi, n1, n2, seed=11, uc=1.5, amp=0.05):
    """Parabola u = uc·(1-(z-1)²) (docs/src/workflow.md:101) + seeded uniform
    noise of amplitude amp·uc on all components."""
    rng = np.random.Generator(np.random.PCG64(seed))
    zc, zi = np.asarray(zc), np.asarray(zi)
    noise = lambda nz: amp * uc * (2 * rng.random((n1, n2, nz)) - 1)
    u = uc * (1 - (zc - 1) ** 2)[None, None, :] + noise(len(zc))
    v = noise(len(zc))
    w = noise(len(zi))
    return np.asfortranarray(u), np.asfortranarray(v), np.asfortranarray(w)


def spectral_noise(model, seed=0, amp=0.25, k0=3.0):
    """Seeded random Fourier coefficients with amplitude amp·exp(−|k|/k0)
    (integer wavenumbers) for every field of `model`, the k1 = 0 column
    Hermitian in k2 (real fields); added to a state it gives a smooth
    broadband perturbation."""
    K1, K2 = model.k1max + 1, 2 * model.k2max + 1
    k1 = np.arange(K1)[:, None]
    k2 = np.concatenate([np.arange(model.k2max + 1), np.arange(-model.k2max, 0)])[None, :]
    scale = amp * np.exp(-np.sqrt(k1 ** 2 + k2 ** 2) / k0)[:, :, None]
    rng = np.random.Generator(np.random.PCG64(seed))
    out = {}
    for f in model.fields:
        n = model.nlayers(f)
        a = (rng.standard_normal((K1, K2, n)) + 1j * rng.standard_normal((K1, K2, n))) * scale
        km = model.k2max
        if km > 0:
            a[0, K2 - km:, :] = np.conj(a[0, 1:km + 1, :][::-1, :])
        a[0, 0, :] = 0
        out[f] = np.asfortranarray(a)
    return out


REG_FFT_A = {16: 4, 24: 6, 32: 8, 48: 8, 64: 8, 72: 9, 96: 12, 128: 16, 144: 12, 192: 16, 256: 16, 384: 24}


def _spread(N, need):
    bh = (N // REG_FFT_A[N]) // 2
    return need + (bh - need) % 16


def _supported(N1, N2, precision):
    N = N2
    return N1 == N2 and N % 2 == 0 and N in REG_FFT_A and N <= (192 if precision == "fp32" else 96)


def c2r_fused(N1, N2, nk1, precision):
    """Whether the engine transforms a space's planes with the fused c2r
    kernel (the engine's own rule): square grid, even length with a register
    FFT, at most 192 (fp32) or 96 (fp64) points, and the plane's intermediate
    (rows wide enough for a row pair's exchange) plus the twiddles within
    160 KiB of LDS."""
    if not _supported(N1, N2, precision):
        return False
    N, A = N2, REG_FFT_A[N2]
    zs = _spread(N, max(nk1, ((N // A) * (A + 1) + 1) // 2))
    elem = 8 if precision == "fp32" else 16
    return (N * zs + N) * elem <= 160 * 1024


def r2c_fused(N1, N2, nk1, precision):
    """The fused r2c: its intermediate, one round of row-pair exchange and the
    twiddles within 160 KiB of LDS, one x2 item per thread (1024 threads)."""
    if not _supported(N1, N2, precision):
        return False
    N, A = N2, REG_FFT_A[N2]
    pairs = 32 if N > 128 else min(N // 2, 64)
    elem = 8 if precision == "fp32" else 16
    lds = (N * _spread(N, nk1) + pairs * (N // A) * (A + 1) + N + nk1) * elem
    return lds <= 160 * 1024 and nk1 * max(A, N // A) <= 1024


def algorithmic_bytes(cfg_sizes, kind, precision, path=None):
    """Minimal HBM bytes of each kernel class per right-hand side over all
    spaces (stage_project: per stage; inputs read once, outputs written once).
    ``cfg_sizes`` holds K1, K2, n3 and the physical sizes of the advection and
    SGS spaces; ``path`` the transform path in use ({"c2r": ..., "r2c": ...}
    as `Model.tuning()` reports it; None: whole-plane wherever a space fits)."""
    bc = 8 if precision == "fp32" else 16
    br = bc // 2
    K1, K2, n3 = cfg_sizes["K1"], cfg_sizes["K2"], cfg_sizes["n3"]
    nk = K1 * K2
    S = nk * (3 * n3 - 1) * bc  # spectral velocity state
    spaces = cfg_sizes["spaces"]  # list of (kind, N1, N2)
    nI, nC = n3 - 1, n3

    def nk1(N1):
        return min(N1 // 2 + 1, K1)

    terms = {"adv": (3 * nC + 3 * nI), "sgs": (6 * nC + 3 * nI)}
    srcs = {"adv": (2 * nC + nI), "sgs": (5 * nC + 3 * nI)}
    c2r_whole = path is None or path["c2r"] != "two-pass"
    r2c_whole = path is None or path["r2c"] != "two-pass"
    fused = [sp for sp in spaces if c2r_whole and c2r_fused(sp[1], sp[2], nk1(sp[1]), precision)]
    paired = [sp for sp in spaces if sp not in fused]
    out = {}
    out["linear"] = 2 * S
    out["c2r_2d"] = sum(terms[k] * (nk * bc + N1 * N2 * br) for k, N1, N2 in fused)
    out["c2r_x2"] = sum(terms[k] * (nk * bc + N2 * nk1(N1) * bc) for k, N1, N2 in paired)
    out["c2r_x1"] = sum(terms[k] * (N2 * nk1(N1) * bc + N1 * N2 * br) for k, N1, N2 in paired)
    fused_r = [sp for sp in spaces if r2c_whole and r2c_fused(sp[1], sp[2], nk1(sp[1]), precision)]
    paired_r = [sp for sp in spaces if sp not in fused_r]
    out["r2c_x1"] = sum(srcs[k] * (N1 * N2 * br + N2 * nk1(N1) * bc) for k, N1, N2 in paired_r)
    # fused r2c: every source plane of a target read (T12 serves two targets),
    # the target read and written
    r2c_2d = 0
    for k, N1, N2 in fused_r:
        per_tgt = [(1, nC), (1, nC), (1, nI)] if k == "adv" else [(3, nC), (3, nC), (3, nI)]
        r2c_2d += sum(ns * L * N1 * N2 * br + 2 * L * nk * bc for ns, L in per_tgt)
    out["r2c_2d"] = r2c_2d
    r2c_x2 = []
    for k, N1, N2 in paired_r:
        per_tgt = [(1, nC), (1, nC), (1, nI)] if k == "adv" else [(3, nC), (3, nC), (3, nI)]
        r2c_x2.append(sum(ns * L * N2 * nk1(N1) * bc + 2 * L * nk * bc for ns, L in per_tgt))
    out["r2c_x2"] = sum(r2c_x2)
    for k, N1, N2 in spaces:
        P = N1 * N2
        if k == "adv":
            out["adv_products"] = P * br * ((3 * nC + 3 * nI) + (2 * nC + nI))
        else:
            # eddy viscosity and stresses in one kernel: 9 terms read, 8 sources
            # written (the C-node viscosity only while the τ33 profile is on)
            out["sgs_stress"] = P * br * ((6 * nC + 3 * nI) + (5 * nC + 3 * nI))
            # learned closure: 9 terms read, 6 stresses written
            out["sgs_mlp"] = P * br * ((6 * nC + 3 * nI) + 6 * nC)
    out["stage_project"] = (3 + 4 + 4) / 3 * S  # stage 1: u,du→utmp; stages 2,3: u,utmp,du→out
    out["check_finite"] = S
    return out
