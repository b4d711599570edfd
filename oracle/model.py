"""Float64 NumPy restatement of the BoundaryLayerDynamics.jl model (test infrastructure only).

Arrays follow the reference layout: spectral fields are complex arrays indexed
``[i1, i2, i3]`` of shape ``(k1max+1, 2*k2max+1, n3)`` with ``i2`` ordered as
``[0..k2max, -k2max..-1]`` (``src/Grids.jl:74-93``); physical fields are real
arrays ``[x1, x2, i3]`` of shape ``(N1, N2, n3)``.  All operations are vectorised
over the vertical layers; the per-layer loops of the reference are restated as
array expressions with the same arithmetic.
"""

import math

import numpy as np
import scipy.fft as sfft

from .ode import ODEProblem, SSPRK33, approxdiv, solve

FFT_WORKERS = -1  # use all host cores for scipy.fft (the CPU baseline is multi-threaded)

# ---------------------------------------------------------------------------
# Domains (src/Domains.jl)
# ---------------------------------------------------------------------------


class SmoothWall:  # Domains.jl:179-185 (docs) / struct at :322
    pass


class RoughWall:  # Domains.jl:331-335
    def __init__(self, roughness, von_karman_constant=0.4):
        self.roughness = roughness
        self.von_karman_constant = von_karman_constant


class FreeSlipBoundary:  # Domains.jl:343
    pass


class CustomBoundary:
    """Domains.jl:360-363.  Values are ``'dirichlet'``, ``'neumann'`` or
    ``('dirichlet', value)`` / ``('neumann', gradient)``."""

    def __init__(self, **behaviors):
        self.behaviors = behaviors


def hasboundarylayer(b):  # Domains.jl:365-367
    if isinstance(b, (RoughWall, SmoothWall)):
        return True
    if isinstance(b, FreeSlipBoundary):
        return False
    raise TypeError("no boundary-layer information for this boundary")


class SinusoidalMapping:  # Domains.jl:234-238
    def __init__(self, eta, variant="auto"):
        assert 0 < eta < 1
        self.parameter = eta
        self.variant = variant


def _instantiate(mapping, l3, bcs):
    """Domains.jl:240-278: (x3(ζ), dx3/dζ(ζ))."""
    if np.isscalar(l3):
        x3min, x3max = 0.0, float(l3)
    else:
        x3min, x3max = float(l3[0]), float(l3[1])
    if mapping is None:  # :275-276
        return (lambda z: x3min + z * (x3max - x3min)), (lambda z: x3max - x3min)
    if isinstance(mapping, tuple):  # :277
        return mapping
    assert x3min < x3max
    eta = float(mapping.parameter)
    l = x3max - x3min
    variant = mapping.variant
    if variant == "auto":
        bls = tuple(hasboundarylayer(b) for b in bcs)
        variant = {(True, True): "symmetric", (True, False): "below", (False, True): "above"}.get(bls)
        if variant is None:
            raise ValueError("Could not determine variant of SinusoidalMapping")
    s = math.sin(eta * math.pi / 2)
    if variant == "symmetric":
        x3 = lambda z: x3min + l / 2 * (1 + math.sin(eta * (2 * z - 1) * math.pi / 2) / s)
        dx3 = lambda z: l * eta * math.pi / 2 / s * math.cos(eta * (2 * z - 1) * math.pi / 2)
    elif variant == "below":
        x3 = lambda z: x3min + l * (1 + math.sin(eta * (z - 1) * math.pi / 2) / s)
        dx3 = lambda z: l * eta * math.pi / 2 / s * math.cos(eta * (z - 1) * math.pi / 2)
    elif variant == "above":
        x3 = lambda z: x3min + l * (math.sin(eta * z * math.pi / 2) / s)
        dx3 = lambda z: l * eta * math.pi / 2 / s * math.cos(eta * z * math.pi / 2)
    else:
        raise ValueError("Unknown variant of SinusoidalMapping")
    return x3, dx3


class Domain:
    """Domains.jl:173-190."""

    def __init__(self, dims, lower_boundary, upper_boundary, mapping=None):
        self.hsize = (float(dims[0]), float(dims[1]))
        l3 = 1.0 if len(dims) == 2 else dims[2]
        self.vmap, self.Dvmap = _instantiate(mapping, l3, (lower_boundary, upper_boundary))
        self.lower_boundary = lower_boundary
        self.upper_boundary = upper_boundary

    def size(self, dim):  # Domains.jl:193-197
        if dim in (1, 2):
            return self.hsize[dim - 1]
        return self.vmap(1.0) - self.vmap(0.0)

    def extrema3(self):  # Domains.jl:205-209
        return self.vmap(0.0), self.vmap(1.0)


# ---------------------------------------------------------------------------
# Grid (src/Grids.jl)
# ---------------------------------------------------------------------------


class Grid:
    """Grids.jl:10-35 on a single process (comm = nothing, proc 1 of 1)."""

    def __init__(self, resolution):
        n1, n2, n3 = resolution
        self.k1max, self.k2max = (n1 - 1) // 2, (n2 - 1) // 2
        if n3 < 3:
            raise ValueError("Due to boundary stencils, the grid needs at least 3 vertical layers")
        self.n3c, self.n3i, self.n3global = n3, n3 - 1, n3
        self.i3min, self.i3max = 1, n3

    def fdsize(self):  # Grids.jl:91
        return (1 + self.k1max, 1 + 2 * self.k2max)

    def nlayers(self, ns):
        return self.n3c if ns == "C" else self.n3i


def wavenumbers(grid):  # Grids.jl:74-83
    k1 = np.arange(0, grid.k1max + 1)
    k2 = np.concatenate([np.arange(0, grid.k2max + 1), np.arange(-grid.k2max, 0)])
    return k1, k2


NODES = {  # Grids.jl:96-123
    "vel1": "C", "vel2": "C", "vel3": "I", "vort1": "I", "vort2": "I", "vort3": "C",
    "vel1_1": "C", "vel1_2": "C", "vel1_3": "I", "vel2_1": "C", "vel2_2": "C", "vel2_3": "I",
    "vel3_1": "I", "vel3_2": "I", "vel3_3": "C", "strain12": "C", "strain13": "I", "strain23": "I",
    "adv1": "C", "adv2": "C", "adv3": "I",
}


def vrange(grid, ns, neighbors=False):
    """Grids.jl:130-141: ζ-values (as floats of the exact rationals)."""
    n = grid.n3global
    if ns == "C":
        if neighbors:  # ζ = 0, 1/N, ..., 1
            return np.array([j / n for j in range(0, n + 1)])
        return np.array([(2 * i - 1) / (2 * n) for i in range(1, n + 1)])
    if neighbors:  # C-nodes around each I-node
        return np.array([(2 * i - 1) / (2 * n) for i in range(1, n + 1)])
    return np.array([i / n for i in range(1, n)])


def dx3factors(domain, grid, ns, neighbors=False):
    """derivatives.jl:10-19: α(ζ) = N3 / (dx3/dζ)."""
    alpha = lambda z: grid.n3global / domain.Dvmap(z)
    z = vrange(grid, ns)
    if neighbors:
        znb = vrange(grid, ns, neighbors=True)
        return np.array([[alpha(znb[i]), alpha(z[i]), alpha(znb[i + 1])] for i in range(len(z))])
    return np.array([alpha(v) for v in z])


def dx1factors(domain, grid):  # Domains.jl:285-286, shape (K1, 1)
    k1, _ = wavenumbers(grid)
    return (1j * (k1 * (2 * np.pi / domain.hsize[0])))[:, None]


def dx2factors(domain, grid):  # Domains.jl:287-288, shape (1, K2)
    _, k2 = wavenumbers(grid)
    return (1j * (k2 * (2 * np.pi / domain.hsize[1])))[None, :]


def second_derivatives(domain, grid, ns):  # derivatives.jl:21-27
    k1, k2 = wavenumbers(grid)
    DD1 = (-(k1 ** 2) * (2 * np.pi / domain.hsize[0]) ** 2)[:, None]
    DD2 = (-(k2 ** 2) * (2 * np.pi / domain.hsize[1]) ** 2)[None, :]
    DD3 = dx3factors(domain, grid, ns, neighbors=True)
    return DD1, DD2, DD3


# ---------------------------------------------------------------------------
# Physical-space sizes and transforms (src/physical_space.jl)
# ---------------------------------------------------------------------------


def _pdsize_quadratic(kmax):  # physical_space.jl:49-56
    if math.log2(1 + kmax).is_integer():
        return 3 * kmax + 3
    return 3 * kmax + (1 if kmax % 2 == 1 else 2)


def pdsize(grid, dealiasing=None):  # physical_space.jl:42-56
    if dealiasing is None:
        return (2 + 2 * grid.k1max, 2 + 2 * grid.k2max)
    if isinstance(dealiasing, tuple):
        return dealiasing
    if dealiasing == "quadratic":
        return (_pdsize_quadratic(grid.k1max), _pdsize_quadratic(grid.k2max))
    raise ValueError(f"Unsupported dealiasing type: {dealiasing}")


def _kmax_pair(n1, n2, K1, K2):
    # physical_space.jl:217-218 / 280-281
    k1m = min(n1 // 2 + 1, K1) - 1
    k2m = (min(n2, K2) - 1) // 2
    return k1m, k2m


def get_field(fd, n1, n2, prefactors=None):
    """physical_space.jl:277-312: zero-padded, unnormalised c2r of every layer.

    ``prefactors`` is a (K1,1), (1,K2) or (K1,K2) array multiplied onto the
    resolved modes before the transform (``get_layer!`` ``prefactors``)."""
    K1, K2, L = fd.shape
    k1m, k2m = _kmax_pair(n1, n2, K1, K2)
    src = fd if prefactors is None else fd * np.asarray(prefactors)[:, :, None]
    buf = np.zeros((n1 // 2 + 1, n2, L), dtype=complex)
    buf[: k1m + 1, : k2m + 1] = src[: k1m + 1, : k2m + 1]
    if k2m > 0:
        buf[: k1m + 1, n2 - k2m:] = src[: k1m + 1, K2 - k2m:]
    return sfft.irfftn(buf, s=(n2, n1), axes=(1, 0), norm="forward", workers=FFT_WORKERS)


def add_field(fd, pd, prefactors=None):
    """physical_space.jl:208-246: fd += prefactor · truncate(rfft(pd) / (N1·N2))."""
    n1, n2, L = pd.shape
    K1, K2, _ = fd.shape
    k1m, k2m = _kmax_pair(n1, n2, K1, K2)
    full = sfft.rfftn(pd, axes=(1, 0), workers=FFT_WORKERS) * (1 / (n1 * n2))
    pre = np.ones((K1, K2)) if prefactors is None else np.broadcast_to(np.asarray(prefactors), (K1, K2))
    fd[: k1m + 1, : k2m + 1] += full[: k1m + 1, : k2m + 1] * pre[: k1m + 1, : k2m + 1, None]
    if k2m > 0:
        fd[: k1m + 1, K2 - k2m:] += full[: k1m + 1, n2 - k2m:] * pre[: k1m + 1, K2 - k2m:, None]
    return fd


def h_range(n):  # physical_space.jl:64-65: ξ = 0, 1/n, ..., (n-1)/n
    return np.array([i / n for i in range(n)])


def _evaluate(fn, x1, x2, x3):
    """fn on the tensor grid; vectorised when fn accepts arrays, else pointwise."""
    X1, X2, X3 = np.meshgrid(x1, x2, x3, indexing="ij")
    try:
        vals = np.asarray(fn(X1, X2, X3), dtype=float)
        if vals.shape == X1.shape:
            return vals
    except (TypeError, ValueError):
        pass
    vals = np.empty(X1.shape)
    for k, z in enumerate(x3):
        for j, y in enumerate(x2):
            for i, x in enumerate(x1):
                vals[i, j, k] = fn(x, y, z)
    return vals


def set_field(fd, fn, domain, grid, ns, n1, n2):
    """physical_space.jl:258-269: evaluate fn on the physical grid, then r2c."""
    x1 = domain.hsize[0] * h_range(n1)
    x2 = domain.hsize[1] * h_range(n2)
    x3 = np.array([domain.vmap(z) for z in vrange(grid, ns)])
    if np.isscalar(fn):
        vals = np.full((n1, n2, len(x3)), float(fn))
    else:
        vals = _evaluate(fn, x1, x2, x3)
    fd[...] = 0
    return add_field(fd, vals)


# ---------------------------------------------------------------------------
# Boundary conditions (src/boundary_conditions.jl)
# ---------------------------------------------------------------------------


class ConstantValue:  # boundary_conditions.jl:36-38
    def __init__(self, value):
        self.value = float(value)


class ConstantGradient:  # boundary_conditions.jl:40-43
    def __init__(self, gradient, difference):
        self.gradient = float(gradient)
        self.difference = float(difference)


def bctype(boundary, field):
    """boundary_conditions.jl:85-95: returns ('dirichlet'|'neumann', value)."""
    if isinstance(boundary, CustomBoundary):
        b = boundary.behaviors[field]
    elif isinstance(boundary, (SmoothWall, RoughWall)):
        b = "dirichlet"
    elif isinstance(boundary, FreeSlipBoundary):
        b = "dirichlet" if field == "vel3" else "neumann"
    else:
        raise TypeError(f"unsupported boundary {boundary!r}")
    if isinstance(b, tuple):
        return b[0], float(b[1])
    return b, 0.0


def init_bcs(field, domain, grid):
    """boundary_conditions.jl:53-79 (the `dx3` for Neumann differences)."""
    out = []
    for boundary, zeta in ((domain.lower_boundary, 0.0), (domain.upper_boundary, 1.0)):
        kind, value = bctype(boundary, field)
        if kind == "dirichlet":
            out.append(ConstantValue(value))
        elif kind == "neumann":
            dx3 = domain.Dvmap(zeta) / grid.n3global
            out.append(ConstantGradient(value, value * dx3))
        else:
            raise ValueError(f"unsupported boundary condition {kind!r}")
    return tuple(out)


def expand_full_C(q, lbc, ubc):
    """boundary_conditions.jl:276-359 on one process: ghost layers below/above
    the C-node field ``q`` (K1,K2,n3) from 4th-order extrapolation.  Spectral
    fields receive the boundary value in mode (0,0) only."""
    l1, l2, l3 = q[:, :, 0], q[:, :, 1], q[:, :, 2]
    if isinstance(lbc, ConstantValue):  # :286-302
        below = (-15 * l1 + 5 * l2 - l3) / 5
        below[0, 0] += 16 * lbc.value / 5
    else:  # ConstantGradient :322-338
        below = (21 * l1 + 3 * l2 - l3) / 23
        below[0, 0] -= 24 * lbc.difference / 23
    u1, u2, u3 = q[:, :, -1], q[:, :, -2], q[:, :, -3]
    if isinstance(ubc, ConstantValue):  # :304-320
        above = (-15 * u1 + 5 * u2 - u3) / 5
        above[0, 0] += 16 * ubc.value / 5
    else:  # :340-356
        above = (21 * u1 + 3 * u2 - u3) / 23
        above[0, 0] -= 24 * (-ubc.difference) / 23
    return below, above


# ---------------------------------------------------------------------------
# Processes (src/processes/*.jl)
# ---------------------------------------------------------------------------


class MomentumAdvection:  # momentum_advection.jl:18-21
    def __init__(self, dealiasing="quadratic"):
        self.dealiasing = dealiasing


class MolecularDiffusion:  # molecular_diffusion.jl:12-18
    def __init__(self, field, diffusivity):
        self.field = field
        self.diffusivity = diffusivity


class Pressure:  # pressure.jl:215-218
    def __init__(self, batch_size=64):
        self.batch_size = batch_size


class ConstantSource:  # sources.jl:13-17
    def __init__(self, field, strength=1.0):
        self.field = field
        self.strength = strength


class ConstantMean:  # sources.jl:47-51
    def __init__(self, field, mean_value=1.0):
        self.field = field
        self.mean_value = mean_value


class StaticSmagorinskyModel:  # static_smagorinsky.jl:15-22
    def __init__(self, dealiasing=None, Cs=0.1, wall_damping=True, wall_damping_exponent=2):
        self.dealiasing = dealiasing
        self.model_constant = Cs
        self.wall_damping = wall_damping
        self.wall_damping_exponent = wall_damping_exponent


def incompressible_flow(viscosity, sgs_model=None, constant_forcing=None, constant_flux=None):
    """BoundaryLayerDynamics.jl:352-368."""
    if constant_flux is not None and constant_forcing is not None:
        raise ValueError("Momentum source set up for both constant forcing and constant flux")
    procs = [MomentumAdvection()] + [MolecularDiffusion(v, viscosity) for v in ("vel1", "vel2", "vel3")]
    procs.append(Pressure())
    if constant_flux is None:
        forcing = 1 if constant_forcing is None else constant_forcing
        forcing = forcing if isinstance(forcing, (tuple, list)) else (forcing,)
        procs += [ConstantSource(v, f) for v, f in zip(("vel1", "vel2"), forcing)]
    else:
        flux = constant_flux if isinstance(constant_flux, (tuple, list)) else (constant_flux,)
        procs += [ConstantMean(v, m) for v, m in zip(("vel1", "vel2"), flux)]
    if sgs_model is not None:
        procs.append(sgs_model)
    return procs


class BatchLDLt:
    """pressure.jl:38-101, 106-201 on one process: batched symmetric Thomas.

    ``dvs`` (nk, nz) diagonals, ``evs`` (nk, nz-1) off-diagonals."""

    def __init__(self, dvs, evs):
        dvs = np.asarray(dvs, dtype=float)
        evs = np.asarray(evs, dtype=float)
        nk, nz = dvs.shape
        self.beta = np.zeros((nk, nz - 1)) if nz > 1 else np.zeros((nk, 0))
        self.beta[:, :] = evs
        g = np.zeros((nk, nz))
        g[:, 0] = 1 / dvs[:, 0]  # :91
        for i in range(1, nz):  # :93-95
            g[:, i] = 1 / (dvs[:, i] - g[:, i - 1] * self.beta[:, i - 1] ** 2)
        self.gamma = g

    def solve(self, b):
        """ldiv! (pressure.jl:167-201) in place on b viewed as (nk, nz)."""
        g, be = self.gamma, self.beta
        nz = g.shape[1]
        b[:, 0] = g[:, 0] * b[:, 0]  # :122-124
        for iz in range(1, nz):  # :126-132
            b[:, iz] = g[:, iz] * (b[:, iz] - be[:, iz - 1] * b[:, iz - 1])
        for iz in range(nz - 2, -1, -1):  # :155-160
            b[:, iz] = b[:, iz] - g[:, iz] * be[:, iz] * b[:, iz + 1]
        return b


def _mixing_length(x3, wall):  # static_smagorinsky.jl:125-127
    if isinstance(wall, FreeSlipBoundary):
        return math.inf
    if isinstance(wall, RoughWall):
        return x3 * wall.von_karman_constant
    raise ValueError("SGS model only supports rough-wall and free-slip boundaries")


class Model:
    """BoundaryLayerDynamics.jl:308-333: grid + discretised processes + state."""

    def __init__(self, resolution, domain, processes):
        self.domain = domain
        self.grid = grid = Grid(resolution)
        flat = []
        for p in processes:
            flat.extend(p if isinstance(p, (tuple, list)) else [p])
        self.processes = flat
        self.fields = []
        for p in flat:  # Processes.jl:326-336 (state_fields, unique, in order)
            fs = ("vel1", "vel2", "vel3") if isinstance(
                p, (MomentumAdvection, Pressure, StaticSmagorinskyModel)) else (p.field,)
            for f in fs:
                if f not in self.fields:
                    self.fields.append(f)
        K1, K2 = grid.fdsize()
        self.state = {f: np.zeros((K1, K2, grid.nlayers(NODES[f])), dtype=complex, order="F")
                      for f in self.fields}
        self._init_processes()

    # -- discretisation (init_process of each process) --------------------
    def _init_processes(self):
        d, g = self.domain, self.grid
        self.D1, self.D2 = dx1factors(d, g), dx2factors(d, g)
        self.alpha_c = dx3factors(d, g, "C")
        self.alpha_i = dx3factors(d, g, "I")
        self.disc = []
        for p in self.processes:
            if isinstance(p, MomentumAdvection):
                dims = pdsize(g, p.dealiasing)
                bcs = {v: init_bcs(v, d, g) for v in ("vel1", "vel2", "vel3")}
                self.disc.append(("adv", dict(dims=dims, bcs=bcs)))
            elif isinstance(p, MolecularDiffusion):
                ns = NODES[p.field]
                self.disc.append(("diff", dict(field=p.field, nu=float(p.diffusivity),
                                               bcs=init_bcs(p.field, d, g),
                                               dd=second_derivatives(d, g, ns), ns=ns)))
            elif isinstance(p, Pressure):
                self.disc.append(("pressure", self._init_pressure()))
            elif isinstance(p, ConstantSource):
                self.disc.append(("source", dict(field=p.field, strength=float(p.strength))))
            elif isinstance(p, ConstantMean):
                if NODES[p.field] != "C":
                    raise ValueError("The constant-mean source currently does not handle I-nodes correctly")
                weight = 1 / self.alpha_c / d.size(3)  # sources.jl:340-344
                self.disc.append(("mean", dict(field=p.field, mean=float(p.mean_value), weight=weight)))
            elif isinstance(p, StaticSmagorinskyModel):
                self.disc.append(("sgs", self._init_sgs(p)))
            else:
                raise TypeError(f"unknown process {p!r}")

    def _init_pressure(self):
        """pressure.jl:231-296."""
        d, g = self.domain, self.grid
        ev = dx3factors(d, g, "I")
        ai_ext = np.array([g.n3global / d.Dvmap(z) for z in vrange(g, "C", neighbors=True)])
        n = g.n3c
        d3g3 = np.zeros(n)
        d3g3[1:] -= ai_ext[1:n]  # dvs[2:end] .-= αi_ext[2:end-1]
        d3g3[: n - 1] -= ai_ext[1:n]  # dvs[1:n3i] .-= αi_ext[2:n3i+1]
        ac = self.alpha_c
        k1, k2 = wavenumbers(g)
        K1, K2 = g.fdsize()
        s1, s2 = 1 / d.hsize[0], 1 / d.hsize[1]
        kk1 = np.tile(k1, K2)  # mode index i1 + K1*i2 (column-major)
        kk2 = np.repeat(k2, K1)
        shift = 4 * np.pi ** 2 * (kk1 ** 2 * s1 ** 2 + kk2 ** 2 * s2 ** 2)
        dvs = d3g3[None, :] - shift[:, None] / ac[None, :]
        dv0 = d3g3.copy()  # k1 = k2 = 0 (adjust_d3g3_diagonal, :287-296)
        last_ev = g.n3global * (1 / d.Dvmap(vrange(g, "C", neighbors=True)[-2]))
        dv0[-1] = -3 * last_ev
        dvs[0] = dv0
        evs = np.broadcast_to(ev, (K1 * K2, n - 1))
        bcs3 = init_bcs("vel3", d, g)
        return dict(solver=BatchLDLt(dvs, evs), bcs3=bcs3)

    def _init_sgs(self, p):
        """static_smagorinsky.jl:44-123."""
        d, g = self.domain, self.grid
        dims = pdsize(g, p.dealiasing)
        D1 = d.size(1) / (2 * g.k1max + 2)
        D2 = d.size(2) / (2 * g.k2max + 2)
        zc, zi = vrange(g, "C"), vrange(g, "I")
        D3c = np.array([d.Dvmap(z) for z in zc]) / g.n3global
        D3i = np.array([d.Dvmap(z) for z in zi]) / g.n3global
        lc = p.model_constant * np.cbrt(D1 * D2 * D3c)
        li = p.model_constant * np.cbrt(D1 * D2 * D3i)
        x3min, x3max = d.vmap(0.0), d.vmap(1.0)
        if p.wall_damping:
            lb, ub = d.lower_boundary, d.upper_boundary
            ml = lambda x3: min(_mixing_length(x3 - x3min, lb), _mixing_length(x3max - x3, ub))
            lcw = np.array([ml(d.vmap(z)) for z in zc])
            liw = np.array([ml(d.vmap(z)) for z in zi])
            N = p.wall_damping_exponent
            lc = (lc ** (-N) + lcw ** (-N)) ** (-1 / N)
            li = (li ** (-N) + liw ** (-N)) ** (-1 / N)

        def init_wall(b, dx3):
            if isinstance(b, RoughWall):
                return ("rough", 1 / (2 * dx3 * math.log(dx3 / b.roughness)),
                        -b.von_karman_constant ** 2 / math.log(dx3 / b.roughness) ** 2)
            if isinstance(b, FreeSlipBoundary):
                return ("freeslip", 0.0, 0.0)
            raise ValueError("SGS model only supports rough-wall and free-slip boundaries")

        dx3_below = d.vmap(zc[0]) - x3min
        dx3_above = x3max - d.vmap(zc[-1])
        lw = init_wall(d.lower_boundary, dx3_below)
        uw = init_wall(d.upper_boundary, dx3_above)
        return dict(dims=dims, lc=lc, li=li, walls=(lw[0], uw[0]),
                    fstrain=(lw[1], -uw[1]), fstress=(lw[2], -uw[2]),
                    bcs3=init_bcs("vel3", d, g), evc=None, evi=None)

    # convenience API mirroring BoundaryLayerDynamics.jl:343-350
    def initialize(self, **ics):
        n1, n2 = pdsize(self.grid)
        for f in ics:
            if f not in self.state:
                raise KeyError(f"Cannot initialize `{f}`: unknown field")
        for f, arr in self.state.items():
            if f in ics:
                set_field(arr, ics[f], self.domain, self.grid, NODES[f], n1, n2)
            else:
                arr[...] = 0

    def __getitem__(self, field):  # State.jl:81-85
        n1, n2 = pdsize(self.grid)
        return get_field(self.state[field], n1, n2)

    def getterm(self, field, ns=None):
        """State.jl:81-99: default-size physical values, optionally interpolated
        from I- to C-nodes with the physical boundary values."""
        vals = self[field]
        if ns is None or ns == NODES[field]:
            return vals
        lbc, ubc = init_bcs(field, self.domain, self.grid)
        if not (isinstance(lbc, ConstantValue) and isinstance(ubc, ConstantValue)):
            raise ValueError("interpolation needs Dirichlet boundary values")
        ext = np.concatenate([np.full_like(vals[:, :, :1], lbc.value), vals,
                              np.full_like(vals[:, :, :1], ubc.value)], axis=2)
        return (ext[:, :, :-1] + ext[:, :, 1:]) / 2

    def coordinates(self, field, dim):  # State.jl:93-95
        n1, n2 = pdsize(self.grid)
        if dim == 1:
            return self.domain.hsize[0] * h_range(n1)
        if dim == 2:
            return self.domain.hsize[1] * h_range(n2)
        return np.array([self.domain.vmap(z) for z in vrange(self.grid, NODES[field])])


# ---------------------------------------------------------------------------
# Rates (Processes.jl:39-62) and projections (Processes.jl:67-75)
# ---------------------------------------------------------------------------


def _add_laplacian(rate, q, pdata, alpha_unused=None):
    """molecular_diffusion.jl:57-72 with the layer expansion of
    boundary_conditions.jl:254-359."""
    lbc, ubc = pdata["bcs"]
    DD1, DD2, DD3 = pdata["dd"]
    nu = pdata["nu"]
    n = q.shape[2]
    if pdata["ns"] == "C":
        below, above = expand_full_C(q, lbc, ubc)
    else:  # I-nodes (:263-274): Dirichlet values are scalars added to every mode
        if not (isinstance(lbc, ConstantValue) and isinstance(ubc, ConstantValue)):
            raise ValueError("diffusion of an I-node field supports Dirichlet boundaries only")
        below = np.full(q.shape[:2], lbc.value, dtype=complex)
        above = np.full(q.shape[:2], ubc.value, dtype=complex)
    ext = np.concatenate([below[:, :, None], q, above[:, :, None]], axis=2)
    a0, a1, a2 = DD3[:, 0], DD3[:, 1], DD3[:, 2]
    lo, mid, hi = ext[:, :, 0:n], ext[:, :, 1:n + 1], ext[:, :, 2:n + 2]
    rate += nu * ((a0 * a1) * lo + ((DD1 + DD2)[:, :, None] - a0 * a1 - a1 * a2) * mid + (a1 * a2) * hi)


def _check_w_bc_adv(bcs):
    for bc in bcs["vel3"]:
        if not isinstance(bc, ConstantValue):
            raise ValueError("advection requires Dirichlet boundary conditions for vel3")
        if bc.value != 0:
            raise ValueError("Advection for non-zero w at boundary not implemented")


def advection_terms(model, state, dims):
    """Physical-space terms of MomentumAdvection (momentum_advection.jl:40-41):
    u, v (C), w (I) and the vorticity computed spectrally (derivatives.jl:199-230,
    always the spectral route, see derivatives.jl:484-506)."""
    n1, n2 = dims
    D1, D2 = model.D1, model.D2
    ai = model.alpha_i
    u_s, v_s, w_s = state["vel1"], state["vel2"], state["vel3"]
    u = get_field(u_s, n1, n2)
    v = get_field(v_s, n1, n2)
    w = get_field(w_s, n1, n2)
    # set_vort1!: D2·w - (-α v¯ + α v⁺);  set_vort2!: (-α u¯ + α u⁺) - D1·w;  set_vort3!: D1 v - D2 u
    o1s = D2[:, :, None] * w_s - (-ai * v_s[:, :, :-1] + ai * v_s[:, :, 1:])
    o2s = (-ai * u_s[:, :, :-1] + ai * u_s[:, :, 1:]) - D1[:, :, None] * w_s
    o3s = D1[:, :, None] * v_s - D2[:, :, None] * u_s
    return u, v, w, get_field(o1s, n1, n2), get_field(o2s, n1, n2), get_field(o3s, n1, n2)


def _advection(model, rates, state, pdata):
    """momentum_advection.jl:44-206."""
    _check_w_bc_adv(pdata["bcs"])
    u, v, w, o1, o2, o3 = advection_terms(model, state, pdata["dims"])
    nc = u.shape[2]
    # products on C-nodes (:180-200): the w-terms at the walls are dropped
    p2w = o2 * w  # I-nodes
    p1w = o1 * w
    below2 = np.concatenate([np.zeros_like(u[:, :, :1]), p2w], axis=2)[:, :, :nc]
    above2 = np.concatenate([p2w, np.zeros_like(u[:, :, :1])], axis=2)
    below1 = np.concatenate([np.zeros_like(u[:, :, :1]), p1w], axis=2)[:, :, :nc]
    above1 = np.concatenate([p1w, np.zeros_like(u[:, :, :1])], axis=2)
    # the zero padding drops the w-term on the wall side of the first/last layer
    adv1 = o3 * v - 0.5 * (below2 + above2)
    adv2 = 0.5 * (below1 + above1) - o3 * u
    # products on I-nodes (:206)
    adv3 = o2 * 0.5 * (u[:, :, :-1] + u[:, :, 1:]) - o1 * 0.5 * (v[:, :, :-1] + v[:, :, 1:])
    add_field(rates["vel1"], adv1)
    add_field(rates["vel2"], adv2)
    add_field(rates["vel3"], adv3)
    return dict(adv1=adv1, adv2=adv2, adv3=adv3)


def _sgs(model, rates, state, pdata):
    """static_smagorinsky.jl:129-249 with the derivative terms of derivatives.jl:104-138,235-241."""
    n1, n2 = pdata["dims"]
    D1, D2 = model.D1, model.D2
    ac, ai = model.alpha_c, model.alpha_i
    u_s, v_s, w_s = state["vel1"], state["vel2"], state["vel3"]
    g = lambda f, pre=None: get_field(f, n1, n2, pre)
    u, v, w = g(u_s), g(v_s), g(w_s)
    u_1, v_2 = g(u_s, D1), g(v_s, D2)
    u_2, v_1 = g(u_s, D2), g(v_s, D1)
    w_1, w_2 = g(w_s, D1), g(w_s, D2)
    lb3, ub3 = pdata["bcs3"]
    for bc in (lb3, ub3):
        if not isinstance(bc, ConstantValue):
            raise ValueError("the SGS model requires Dirichlet boundary conditions for vel3")
    # vel3_3 on C (dx3_i2c!, derivatives.jl:449-476; physical ConstantValue)
    wext = np.concatenate([np.full_like(u[:, :, :1], lb3.value), w, np.full_like(u[:, :, :1], ub3.value)], axis=2)
    w_3 = -ac * wext[:, :, :-1] + ac * wext[:, :, 1:]
    # vel1_3, vel2_3 on I (dx3_c2i!, derivatives.jl:442-448,459)
    u_3 = -ai * u[:, :, :-1] + ai * u[:, :, 1:]
    v_3 = -ai * v[:, :, :-1] + ai * v[:, :, 1:]
    S11, S22, S33 = u_1, v_2, w_3
    S12 = (u_2 + v_1) / 2
    S13 = (u_3 + w_1) / 2
    S23 = (v_3 + w_2) / 2
    # total_strain! (:202-233)
    sc = S11 ** 2 + S22 ** 2 + S33 ** 2 + 2 * S12 ** 2
    si = 2 * S13 ** 2 + 2 * S23 ** 2
    for S, cnt in ((S11, 1), (S22, 1), (S33, 1), (S12, 2)):
        si += cnt * (S[:, :, :-1] + S[:, :, 1:]) ** 2 / 4
    lw, uw = pdata["walls"]
    fs_l, fs_u = pdata["fstrain"]
    for S, vel in ((S13, u), (S23, v)):
        # interior C-layers: interpolate I-node values; wall layers use the wall model
        sc[:, :, 1:-1] += 2 * (S[:, :, :-1] + S[:, :, 1:]) ** 2 / 4
        if lw == "rough":
            sc[:, :, 0] += 2 * (vel[:, :, 0] * fs_l) ** 2
        else:
            sc[:, :, 0] += 2 * (0.0 + S[:, :, 0]) ** 2 / 4
        if uw == "rough":
            sc[:, :, -1] += 2 * (vel[:, :, -1] * fs_u) ** 2
        else:
            sc[:, :, -1] += 2 * (S[:, :, -1] + 0.0) ** 2 / 4
    sc = np.sqrt(2 * sc)
    si = np.sqrt(2 * si)
    evc = sc * pdata["lc"] ** 2
    evi = si * pdata["li"] ** 2
    pdata["evc"], pdata["evi"] = evc, evi
    # stresses and their (partial) divergences (:145-186)
    t11 = -2 * evc * S11
    add_field(rates["vel1"], -t11, D1)
    t22 = -2 * evc * S22
    add_field(rates["vel2"], -t22, D2)
    t33 = -2 * evc * S33
    add_field(rates["vel3"], -(-ai * t33[:, :, :-1] + ai * t33[:, :, 1:]))
    t12 = -2 * evc * S12
    add_field(rates["vel1"], -t12, D2)
    add_field(rates["vel2"], -t12, D1)
    fst_l, fst_u = pdata["fstress"]
    mag_l = np.sqrt(u[:, :, 0] ** 2 + v[:, :, 0] ** 2)
    mag_u = np.sqrt(u[:, :, -1] ** 2 + v[:, :, -1] ** 2)
    out = dict(evc=evc, evi=evi, sgs11=t11, sgs22=t22, sgs33=t33, sgs12=t12)
    for t_dir, Sx, vel, rate_f, rate_h, pre in ((1, S13, u, "vel1", "vel3", D1), (2, S23, v, "vel2", "vel3", D2)):
        ti = -2 * evi * Sx
        add_field(rates[rate_h], -ti, pre)
        # wall stress (stress_bc, :195-200); free slip = ConstantValue(0)
        bl = fst_l * vel[:, :, 0] * mag_l if lw == "rough" else np.zeros_like(mag_l)
        bu = fst_u * vel[:, :, -1] * mag_u if uw == "rough" else np.zeros_like(mag_u)
        ext = np.concatenate([bl[:, :, None], ti, bu[:, :, None]], axis=2)
        div = -ac * ext[:, :, :-1] + ac * ext[:, :, 1:]
        add_field(rates[rate_f], -div)
        out[f"sgs{t_dir}3"] = ext  # I-node stress with the wall stresses as boundary layers
    return out


def compute_rates(model, rates, state, terms=None):
    """Processes.jl:39-62: reset, linear (spectral) terms, nonlinear (physical) terms.
    ``terms`` (a dict) receives the physical-space terms that the reference logs
    (log_sample!: adv1-3, sgs11-33)."""
    for r in rates.values():
        r[...] = 0
    for kind, pd in model.disc:  # linear processes first
        if kind == "diff":
            _add_laplacian(rates[pd["field"]], state[pd["field"]], pd)
        elif kind == "source":
            rates[pd["field"]][0, 0, :] += pd["strength"]  # sources.jl:32-35
    for kind, pd in model.disc:
        if kind == "adv":
            r = _advection(model, rates, state, pd)
        elif kind == "sgs":
            r = _sgs(model, rates, state, pd)
        else:
            continue
        if terms is not None:
            terms.update(r)
    return rates


def _project_pressure(model, state, pdata):
    """pressure.jl:300-426."""
    lbc3, ubc3 = pdata["bcs3"]
    if not (isinstance(lbc3, ConstantValue) and isinstance(ubc3, ConstantValue)):
        raise ValueError("The pressure solver only supports Dirichlet BCs")
    u, v, w = state["vel1"], state["vel2"], state["vel3"]
    D1, D2 = model.D1[:, :, None], model.D2[:, :, None]
    hf = 1 / model.alpha_c
    K1, K2, n = u.shape
    w_lo = np.concatenate([np.zeros((K1, K2, 1), complex), w], axis=2)
    w_hi = np.concatenate([w, np.zeros((K1, K2, 1), complex)], axis=2)
    p = u * D1 * hf + v * D2 * hf - w_lo + w_hi  # div! (:368-378)
    p[0, 0, 0] -= lbc3.value
    p[0, 0, -1] += ubc3.value
    if lbc3.value != ubc3.value:
        raise ValueError("The mean flow must be equal at both vertical boundaries.")
    p[0, 0, -1] = 0  # adjust_pressure_rhs! (:396)
    pf = p.reshape((K1 * K2, n), order="F")
    pf = pdata["solver"].solve(pf)
    p = pf.reshape((K1, K2, n), order="F")
    ai = model.alpha_i
    u += -1 * D1 * p  # add_gradient! (:405-426)
    v += -1 * D2 * p
    w += -1 * (-ai * p[:, :, :-1] + ai * p[:, :, 1:])
    return p


def apply_projections(model, state):
    """Processes.jl:67-75."""
    for kind, pd in model.disc:
        if kind == "pressure":
            _project_pressure(model, state, pd)
        elif kind == "mean":  # sources.jl:71-77
            f = state[pd["field"]]
            cur = np.sum(f[0, 0, :].real * pd["weight"])
            f[0, 0, :] += pd["mean"] - cur
    return state


def rate_fn(model):
    """BoundaryLayerDynamics.jl:149-158 (the `rate!` closure)."""
    def rate(du, u, t, checkpoint=False):
        rates = dict(zip(model.fields, du))
        state = dict(zip(model.fields, u))
        compute_rates(model, rates, state)
        for r in du:
            if not np.all(np.isfinite(r)):
                raise FloatingPointError("The simulation has diverged.")
    return rate


def projection_fn(model):  # BoundaryLayerDynamics.jl:161-165
    def projection(u):
        apply_projections(model, dict(zip(model.fields, u)))
    return projection


def evolve(model, tspan, dt, method=None):
    """BoundaryLayerDynamics.jl:190-219 without output modules."""
    method = SSPRK33() if method is None else method
    if np.isscalar(tspan):
        t1, t2 = 0.0, float(tspan)
    else:
        t1, t2 = float(tspan[0]), float(tspan[-1])
    u0 = [model.state[f] for f in model.fields]
    prob = ODEProblem(rate_fn(model), u0, (t1, t2), projection=projection_fn(model), checkpoint=True)
    if t2 > t1:
        nt = approxdiv(t2 - t1, dt)
        cps = np.linspace(t1 + dt, t2, nt)
        solve(prob, method, dt, checkpoints=cps)
    return prob
