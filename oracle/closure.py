"""CPU oracle of the learned SGS closure and of its filtered-DNS sampler
(test infrastructure only, float64 NumPy).

The closure has no counterpart in the reference.  It takes the place of the
stresses of the static Smagorinsky model (static_smagorinsky.jl:129-249, the
oracle's ``model._sgs``) and keeps that model's wall stress
(static_smagorinsky.jl:195-200) and the structure of its rate terms
(static_smagorinsky.jl:145-186): τ11, τ22, τ12 at C-nodes, τ33 differenced to
the I-nodes, τ13 and τ23 on the I-nodes differenced back to the C-nodes.

Definitions restated here (the GPU kernels follow the same ones):

* velocity gradient g_ij = ∂u_i/∂x_j at each C-node of the default physical
  grid: g11, g12, g21, g22 spectral; g33 = α_c (w⁺ − w⁻) with the boundary
  value of w at the walls; g13, g23 the mean of α_i (u_{l+1} − u_l) over the
  I-layers that exist below and above; g31, g32 the mean of the I-node values
  below and above with 0 at a boundary;
* features x = (g/|g| row-major, wall feature of the layer);
* stress τ = Δ²|g|² y(x) with the trace of the normal components removed; the
  Smagorinsky stress −2 l_c² |S| S of the C-node where y is not finite or |g|
  vanishes, or (with the dissipation guard) where −τ:S < 0;
* samples: box filter of F×F cells of each C-layer, τ_ij = avg(u_i u_j) −
  avg(u_i) avg(u_j) with w interpolated to the C-nodes, target τ^dev/(Δ²|ḡ|²).

Pins (tests/test_oracle_closure.py): a network with zero output reduces the
rates to the Smagorinsky model's at a vanishing model constant (only the wall
stress remains); a constant-output network on a shear flow u = U sin(k x2)
gives the closed-form rates; the sampler is checked against explicit loops
over analytic point values; the gradient definition against analytic
derivatives of trigonometric fields.
"""

import numpy as np

from . import mlp as M
from .model import ConstantValue, add_field, get_field


def default_size(model):
    g = model.grid
    return 2 * g.k1max + 2, 2 * g.k2max + 2


def physical_terms(model, state, n1=None, n2=None):
    """u, v (C), w (I), horizontal derivatives (physical_space.jl:277-312 with
    the derivative prefactors of derivatives.jl:404-418) on an n1 × n2 grid."""
    if n1 is None:
        n1, n2 = default_size(model)
    g = lambda f, pre=None: get_field(f, n1, n2, pre)
    us, vs, ws = state["vel1"], state["vel2"], state["vel3"]
    D1, D2 = model.D1, model.D2
    return dict(u=g(us), v=g(vs), w=g(ws), u_1=g(us, D1), u_2=g(us, D2), v_1=g(vs, D1), v_2=g(vs, D2),
                w_1=g(ws, D1), w_2=g(ws, D2))


def w_boundary(model):
    """Physical boundary values of vel3 (Dirichlet, as the SGS model requires)."""
    from .model import init_bcs
    lb, ub = init_bcs("vel3", model.domain, model.grid)
    for bc in (lb, ub):
        if not isinstance(bc, ConstantValue):
            raise ValueError("the closure requires Dirichlet boundary conditions for vel3")
    return lb.value, ub.value


def velocity_gradient(model, t):
    """g [n1, n2, n3, 9] (row-major ij) at the C-nodes from physical_terms `t`."""
    ac, ai = model.alpha_c, model.alpha_i
    u, v, w = t["u"], t["v"], t["w"]
    n3 = u.shape[2]
    wl, wu = w_boundary(model)
    wext = np.concatenate([np.full_like(u[:, :, :1], wl), w, np.full_like(u[:, :, :1], wu)], axis=2)
    g33 = ac * (wext[:, :, 1:] - wext[:, :, :-1])
    u3 = ai * (u[:, :, 1:] - u[:, :, :-1])  # interior I-layers 0..n3-2
    v3 = ai * (v[:, :, 1:] - v[:, :, :-1])

    def c_mean(q):  # mean over the I-layers that exist below/above each C-layer
        out = np.zeros_like(u)
        cnt = np.zeros(n3)
        out[:, :, 1:] += q
        cnt[1:] += 1
        out[:, :, :-1] += q
        cnt[:-1] += 1
        return out / np.maximum(cnt, 1)

    def c_half(q):  # (below + above)/2 with 0 at the boundaries
        out = np.zeros_like(u)
        out[:, :, 1:] += q
        out[:, :, :-1] += q
        return out / 2

    g = np.stack([t["u_1"], t["u_2"], c_mean(u3), t["v_1"], t["v_2"], c_mean(v3),
                  c_half(t["w_1"]), c_half(t["w_2"]), g33], axis=3)
    return g


def sample(model, state, F, delta, xw):
    """Box-filtered samples: (X [n, 16], Y [n, 16]) in the order
    c = i1c + N1c·(i2c + N2c·l)."""
    t = physical_terms(model, state)
    g = velocity_gradient(model, t)
    u, v, w = t["u"], t["v"], t["w"]
    n1, n2, n3 = u.shape
    wl, wu = w_boundary(model)
    wext = np.concatenate([np.full_like(u[:, :, :1], wl), w, np.full_like(u[:, :, :1], wu)], axis=2)
    wc = (wext[:, :, 1:] + wext[:, :, :-1]) / 2
    n1c, n2c = n1 // F, n2 // F

    def box(q):  # [n1, n2, n3, ...] → [n1c, n2c, n3, ...]
        sh = q.shape
        r = q.reshape(n1c, F, n2c, F, *sh[2:])
        return r.mean(axis=(1, 3))

    G = box(g)
    U = [box(a) for a in (u, v, wc)]
    prods = [(u, u), (v, v), (wc, wc), (u, v), (u, wc), (v, wc)]
    idx = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
    tau = np.stack([box(a * b) - U[i] * U[j] for (a, b), (i, j) in zip(prods, idx)], axis=3)
    tr = (tau[..., 0] + tau[..., 1] + tau[..., 2]) / 3
    for m in range(3):
        tau[..., m] -= tr
    gn2 = np.sum(G * G, axis=3)
    sc = delta ** 2 * gn2
    ig = np.where(gn2 > 0, 1 / np.sqrt(np.where(gn2 > 0, gn2, 1)), 0)
    n = n1c * n2c * n3
    X = np.zeros((n, 16))
    Y = np.zeros((n, 16))
    flat = lambda a: a.reshape(n, *a.shape[3:], order="F") if a.ndim > 3 else a.reshape(n, order="F")
    X[:, :9] = flat(G) * flat(ig)[:, None]
    X[:, 9] = np.repeat(np.asarray(xw), n1c * n2c)
    scf = flat(sc)
    Y[:, :6] = np.where(scf[:, None] > 0, flat(tau) / np.where(scf > 0, scf, 1)[:, None], 0)
    Y[:, 15] = scf
    return X, Y


def closure_stress(model, t, lc, delta, xw, y_fn, guard=False):
    """τ [n1, n2, n3, 6] (11, 22, 33, 12, 13, 23) at the C-nodes; y_fn maps the
    features [n, 16] to the normalised outputs [n, 6].  With `guard`, points
    where the learned stress is not dissipative (−τ:S < 0) take the
    Smagorinsky stress."""
    g = velocity_gradient(model, t)
    n1, n2, n3, _ = g.shape
    gn2 = np.sum(g * g, axis=3)
    ig = np.where(gn2 > 0, 1 / np.sqrt(np.where(gn2 > 0, gn2, 1)), 0)
    X = np.zeros((n1, n2, n3, 16))
    X[..., :9] = g * ig[..., None]
    X[..., 9] = np.asarray(xw)[None, None, :]
    y = np.asarray(y_fn(X.reshape(-1, 16)), dtype=np.float64).reshape(n1, n2, n3, 6)
    tau = (delta ** 2 * gn2)[..., None] * y
    ok = np.all(np.isfinite(tau), axis=3) & (gn2 > 0)
    tr = (tau[..., 0] + tau[..., 1] + tau[..., 2]) / 3
    for m in range(3):
        tau[..., m] -= tr
    # Smagorinsky stress of the C-node where the network output is not usable
    S11, S22, S33 = g[..., 0], g[..., 4], g[..., 8]
    S12, S13, S23 = (g[..., 1] + g[..., 3]) / 2, (g[..., 2] + g[..., 6]) / 2, (g[..., 5] + g[..., 7]) / 2
    if guard:
        tz = np.nan_to_num(tau)
        eps = -(tz[..., 0] * S11 + tz[..., 1] * S22 + tz[..., 2] * S33
                + 2 * (tz[..., 3] * S12 + tz[..., 4] * S13 + tz[..., 5] * S23))
        ok &= eps >= 0
    mag = np.sqrt(2 * (S11 ** 2 + S22 ** 2 + S33 ** 2 + 2 * (S12 ** 2 + S13 ** 2 + S23 ** 2)))
    nu = np.asarray(lc)[None, None, :] ** 2 * mag
    smag = np.stack([-2 * nu * S for S in (S11, S22, S33, S12, S13, S23)], axis=3)
    return np.where(ok[..., None], np.nan_to_num(tau), smag)


def network(params):
    """y_fn of the bf16-operand network of oracle.mlp."""
    return lambda X: M.forward(params, X)[0][:, :6]


def add_closure_rates(model, rates, state, pdata, delta, xw, y_fn, guard=False):
    """Rates of the learned stresses on the SGS grid of `pdata` (the oracle's
    Smagorinsky discretisation of the fallback model), assembled as
    static_smagorinsky.jl:145-186."""
    n1, n2 = pdata["dims"]
    D1, D2 = model.D1, model.D2
    ac, ai = model.alpha_c, model.alpha_i
    t = physical_terms(model, state, n1, n2)
    tau = closure_stress(model, t, pdata["lc"], delta, xw, y_fn, guard)
    t11, t22, t33, t12, t13, t23 = (tau[..., m] for m in range(6))
    add_field(rates["vel1"], -t11, D1)
    add_field(rates["vel2"], -t22, D2)
    add_field(rates["vel3"], -(ai * (t33[:, :, 1:] - t33[:, :, :-1])))
    add_field(rates["vel1"], -t12, D2)
    add_field(rates["vel2"], -t12, D1)
    u, v = t["u"], t["v"]
    lw, uw = pdata["walls"]
    fst_l, fst_u = pdata["fstress"]
    mag_l = np.sqrt(u[:, :, 0] ** 2 + v[:, :, 0] ** 2)
    mag_u = np.sqrt(u[:, :, -1] ** 2 + v[:, :, -1] ** 2)
    for tc, vel, rate_f, pre in ((t13, u, "vel1", D1), (t23, v, "vel2", D2)):
        ti = (tc[:, :, :-1] + tc[:, :, 1:]) / 2
        add_field(rates["vel3"], -ti, pre)
        bl = fst_l * vel[:, :, 0] * mag_l if lw == "rough" else np.zeros_like(mag_l)
        bu = fst_u * vel[:, :, -1] * mag_u if uw == "rough" else np.zeros_like(mag_u)
        ext = np.concatenate([bl[:, :, None], ti, bu[:, :, None]], axis=2)
        add_field(rates[rate_f], -(ac * (ext[:, :, 1:] - ext[:, :, :-1])))
    return rates


def closure_rates(model, state, delta, xw, y_fn, guard=False):
    """compute_rates! (Processes.jl:39-62) of a model whose SGS process is the
    learned closure with `model`'s Smagorinsky discretisation as fallback."""
    from .model import compute_rates
    sgs = [pd for kind, pd in model.disc if kind == "sgs"]
    if len(sgs) != 1:
        raise ValueError("the model needs exactly one SGS process (the fallback)")
    disc = model.disc
    model.disc = [(k, pd) for k, pd in disc if k != "sgs"]
    try:
        rates = {f: np.zeros_like(state[f]) for f in state}
        compute_rates(model, rates, state)
    finally:
        model.disc = disc
    return add_closure_rates(model, rates, state, sgs[0], delta, xw, y_fn, guard)


def wall_feature(model, delta):
    """1/(1 + d/Δ) per C-layer, d the distance to the nearest boundary that is
    not free-slip (0 without one)."""
    from .model import FreeSlipBoundary, vrange
    d = model.domain
    zc = np.array([d.vmap(z) for z in vrange(model.grid, "C")])
    x3min, x3max = d.vmap(0.0), d.vmap(1.0)
    dist = np.full(zc.shape, np.inf)
    if not isinstance(d.lower_boundary, FreeSlipBoundary):
        dist = np.minimum(dist, zc - x3min)
    if not isinstance(d.upper_boundary, FreeSlipBoundary):
        dist = np.minimum(dist, x3max - zc)
    return np.where(np.isinf(dist), 0.0, 1 / (1 + dist / delta))
