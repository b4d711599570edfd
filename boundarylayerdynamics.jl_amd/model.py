"""Host-side mirror of the BoundaryLayerDynamics.jl model API over the HIP engine.

The reference builds a `Model` from a resolution, a `Domain` and a process list
(src/BoundaryLayerDynamics.jl:72-97) and advances it with `evolve!` (:190-219).
This module keeps those names and meanings; the discretisation constants that
the reference's `init_process` methods compute are computed here (float64, on
the host, once) and handed to the engine through `bld_config`.  Everything that
runs per time step runs on the GPU.

Arrays use the reference layout: spectral fields are complex128 Fortran-ordered
arrays of shape (k1max+1, 2*k2max+1, n3) (vel3: n3-1 layers), physical fields
are float64 Fortran-ordered (N1, N2, layers).
"""

import ctypes
import math
import os
import warnings

import numpy as np

from . import _lib
from ._lib import BldConfig, BldError, dptr, fptr

# ---------------------------------------------------------------------------
# Domain and boundaries (src/Domains.jl)
# ---------------------------------------------------------------------------


class SmoothWall:
    """No-slip, no-penetration wall (Domains.jl:322)."""


class RoughWall:
    """Rough wall with roughness length z0 (Domains.jl:331-335)."""

    def __init__(self, roughness, von_karman_constant=0.4):
        self.roughness = roughness
        self.von_karman_constant = von_karman_constant


class FreeSlipBoundary:
    """No penetration, vanishing tangential gradients (Domains.jl:343)."""


class CustomBoundary:
    """Explicit boundary conditions per field (Domains.jl:360-363): values are
    'dirichlet', 'neumann', ('dirichlet', value) or ('neumann', gradient)."""

    def __init__(self, **behaviors):
        self.behaviors = behaviors


class SinusoidalMapping:
    """Grid stretching x3 ∝ sin(ζηπ/2)/sin(ηπ/2) (Domains.jl:234-273)."""

    def __init__(self, eta, variant="auto"):
        if not 0 < eta < 1:
            raise ValueError("SinusoidalMapping requires 0 < η < 1")
        self.parameter = eta
        self.variant = variant


def _has_boundary_layer(b):  # Domains.jl:365-367
    if isinstance(b, (RoughWall, SmoothWall)):
        return True
    if isinstance(b, FreeSlipBoundary):
        return False
    raise ValueError("Could not determine variant of SinusoidalMapping")


class Domain:
    """Domains.jl:173-190: periodic in x1, x2; bounded in x3 (optionally mapped)."""

    def __init__(self, dims, lower_boundary, upper_boundary, mapping=None):
        self.hsize = (float(dims[0]), float(dims[1]))
        l3 = 1.0 if len(dims) == 2 else dims[2]
        x3min, x3max = (0.0, float(l3)) if np.isscalar(l3) else (float(l3[0]), float(l3[1]))
        self.lower_boundary, self.upper_boundary = lower_boundary, upper_boundary
        if mapping is None:
            self.vmap = lambda z: x3min + z * (x3max - x3min)
            self.Dvmap = lambda z: x3max - x3min
        elif isinstance(mapping, tuple):
            self.vmap, self.Dvmap = mapping
        else:
            eta = float(mapping.parameter)
            l = x3max - x3min
            variant = mapping.variant
            if variant == "auto":
                bls = (_has_boundary_layer(lower_boundary), _has_boundary_layer(upper_boundary))
                variant = {(True, True): "symmetric", (True, False): "below", (False, True): "above"}.get(bls)
                if variant is None:
                    raise ValueError("Could not determine variant of SinusoidalMapping")
            s = math.sin(eta * math.pi / 2)
            shift = {"symmetric": lambda z: 2 * z - 1, "below": lambda z: z - 1, "above": lambda z: z}[variant]
            if variant == "symmetric":
                self.vmap = lambda z: x3min + l / 2 * (1 + math.sin(eta * shift(z) * math.pi / 2) / s)
                self.Dvmap = lambda z: l * eta * math.pi / 2 / s * math.cos(eta * shift(z) * math.pi / 2)
            elif variant == "below":
                self.vmap = lambda z: x3min + l * (1 + math.sin(eta * shift(z) * math.pi / 2) / s)
                self.Dvmap = lambda z: l * eta * math.pi / 2 / s * math.cos(eta * shift(z) * math.pi / 2)
            else:
                self.vmap = lambda z: x3min + l * (math.sin(eta * shift(z) * math.pi / 2) / s)
                self.Dvmap = lambda z: l * eta * math.pi / 2 / s * math.cos(eta * shift(z) * math.pi / 2)

    def size(self, dim):
        return self.hsize[dim - 1] if dim in (1, 2) else self.vmap(1.0) - self.vmap(0.0)


# ---------------------------------------------------------------------------
# Processes (src/processes/*.jl) and ODE methods (src/ODEMethods.jl)
# ---------------------------------------------------------------------------


class MomentumAdvection:
    def __init__(self, dealiasing="quadratic"):
        self.dealiasing = dealiasing


class MolecularDiffusion:
    def __init__(self, field, diffusivity):
        self.field = field
        self.diffusivity = diffusivity


class Pressure:
    def __init__(self, batch_size=64):  # batch_size only pipelines MPI ranks in the reference
        self.batch_size = batch_size


class ConstantSource:
    def __init__(self, field, strength=1.0):
        self.field = field
        self.strength = strength


class ConstantMean:
    """sources.jl:47-77: a projection that keeps the z-weighted mean of the
    (0,0) mode of `field` at `mean_value` (constant flux)."""

    def __init__(self, field, mean_value=1.0):
        self.field = field
        self.mean_value = mean_value


class StaticSmagorinskyModel:
    def __init__(self, dealiasing=None, Cs=0.1, wall_damping=True, wall_damping_exponent=2):
        self.dealiasing = dealiasing
        self.model_constant = Cs
        self.wall_damping = wall_damping
        self.wall_damping_exponent = wall_damping_exponent


class LearnedSGSModel:
    """Learned SGS closure (no reference counterpart), a process next to
    StaticSmagorinskyModel (static_smagorinsky.jl:39-42) with the same state
    fields and rate terms: the network of `closure.MLP` gives the deviatoric
    stress at every C-node from the normalised velocity gradient and a wall
    feature; `fallback`, a StaticSmagorinskyModel, supplies the wall model and
    the eddy-viscosity stress wherever the network output is not finite.

    ``params``: flat float32 parameters (closure.pack layout) or a dict;
    ``delta``: Δ of the stress scale Δ²|∇u|² (default: the geometric mean of the
    horizontal grid spacings of the default physical grid); ``guard``: points
    where the learned stress is not dissipative (−τ:S < 0) take the fallback's
    stress."""

    def __init__(self, params, fallback=None, delta=None, guard=True):
        self.params = params
        self.guard = guard
        self.fallback = StaticSmagorinskyModel() if fallback is None else fallback
        self.delta = delta
        self.dealiasing = self.fallback.dealiasing


def wall_feature(domain, n3, delta):
    """Network input 10 per C-layer: 1/(1 + d/Δ) with d the distance to the
    nearest wall that is not free-slip (0 without such a wall)."""
    zc = [domain.vmap(z) for z in _zeta(n3, "C")]
    x3min, x3max = domain.vmap(0.0), domain.vmap(1.0)
    out = []
    for z in zc:
        d = math.inf
        if not isinstance(domain.lower_boundary, FreeSlipBoundary):
            d = min(d, z - x3min)
        if not isinstance(domain.upper_boundary, FreeSlipBoundary):
            d = min(d, x3max - z)
        out.append(0.0 if math.isinf(d) else 1.0 / (1.0 + d / delta))
    return np.array(out, dtype=np.float64)


class SSPRK33:
    """Three-stage SSP Runge-Kutta (ODEMethods.jl:50,147-166), the default method."""

    code = _lib.BLD_SSPRK33


class SSPRK22:
    """Two-stage SSP Runge-Kutta (ODEMethods.jl:49,133-145)."""

    code = _lib.BLD_SSPRK22


class Euler:
    """Forward Euler (ODEMethods.jl:47,100-103)."""

    code = _lib.BLD_EULER


class AB2:
    """Two-step Adams–Bashforth with a Runge-Kutta first step
    (ODEMethods.jl:48,105-131); dt must stay constant."""

    code = _lib.BLD_AB2


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


def closedchannelflow(Re, dims, precision="fp64", **kw):  # BoundaryLayerDynamics.jl:134-138
    dom = Domain((4 * math.pi, 2 * math.pi, 2), SmoothWall(), SmoothWall())
    return Model(dims, dom, incompressible_flow(1 / Re, **kw), precision=precision)


def openchannelflow(Re, dims, roughness_length=None, precision="fp64", **kw):  # :140-145
    wall = SmoothWall() if roughness_length is None else RoughWall(roughness_length)
    dom = Domain((4 * math.pi, 2 * math.pi, 1), wall, FreeSlipBoundary())
    return Model(dims, dom, incompressible_flow(1 / Re, **kw), precision=precision)


# ---------------------------------------------------------------------------
# grid helpers (src/Grids.jl, src/physical_space.jl)
# ---------------------------------------------------------------------------

NODES = {"vel1": "C", "vel2": "C", "vel3": "I"}
FIELD_INDEX = {"vel1": 0, "vel2": 1, "vel3": 2}


def pdsize(kmax_pair, dealiasing):
    """physical_space.jl:42-56."""
    if dealiasing is None:
        return tuple(2 + 2 * k for k in kmax_pair)
    if isinstance(dealiasing, tuple):
        return dealiasing
    if dealiasing == "quadratic":
        return tuple(3 * k + 3 if math.log2(1 + k).is_integer() else 3 * k + (1 if k % 2 else 2)
                     for k in kmax_pair)
    raise ValueError(f"Unsupported dealiasing type: {dealiasing}")


def _zeta(n3, ns, neighbors=False):  # Grids.jl:130-141
    if ns == "C":
        return [j / n3 for j in range(n3 + 1)] if neighbors else [(2 * i - 1) / (2 * n3) for i in range(1, n3 + 1)]
    return [(2 * i - 1) / (2 * n3) for i in range(1, n3 + 1)] if neighbors else [i / n3 for i in range(1, n3)]


def _bctype(boundary, field):  # boundary_conditions.jl:85-95
    if isinstance(boundary, CustomBoundary):
        b = boundary.behaviors[field]
    elif isinstance(boundary, (SmoothWall, RoughWall)):
        b = "dirichlet"
    elif isinstance(boundary, FreeSlipBoundary):
        b = "dirichlet" if field == "vel3" else "neumann"
    else:
        raise ValueError(f"unsupported boundary {boundary!r}")
    kind, value = (b[0], float(b[1])) if isinstance(b, tuple) else (b, 0.0)
    if kind not in ("dirichlet", "neumann"):
        raise ValueError(f"unsupported boundary condition {kind!r}")
    return kind, value


# ---------------------------------------------------------------------------
# Model
# ---------------------------------------------------------------------------


class Discretization:
    """The constants that the reference's init_process methods compute
    (src/processes/*.jl), packed into the engine's `bld_config`."""

    def __init__(self, resolution, domain, processes, precision="fp64"):
        n1, n2, n3 = resolution
        self.k1max, self.k2max = (n1 - 1) // 2, (n2 - 1) // 2
        if n3 < 3:
            raise ValueError("Due to boundary stencils, the grid needs at least 3 vertical layers")
        if precision not in ("fp64", "fp32"):
            raise ValueError("precision must be 'fp64' or 'fp32'")
        self.n3 = n3
        self.domain = domain
        flat = []
        for p in processes:
            flat.extend(p if isinstance(p, (tuple, list)) else [p])
        self.processes = flat
        self.precision = precision
        self.fields = []
        for p in flat:  # Processes.jl:326-336
            fs = ("vel1", "vel2", "vel3") if isinstance(
                p, (MomentumAdvection, Pressure, StaticSmagorinskyModel, LearnedSGSModel)) else (p.field,)
            for f in fs:
                if f not in FIELD_INDEX:
                    raise NotImplementedError(f"field {f!r}: the GPU engine advances vel1, vel2, vel3")
                if f not in self.fields:
                    self.fields.append(f)
        self._keep = []
        self.closure = None
        self.config = self._config()

    def closure_args(self):
        """(params float32, Δ, wall feature, flags) of the learned closure, or None."""
        p = self.closure
        if p is None:
            return None
        from . import closure as _cl
        flat = np.ascontiguousarray(p.params if isinstance(p.params, np.ndarray) else _cl.pack(p.params),
                                    dtype=np.float32)
        if flat.shape != (_cl.NPARAM,):
            raise ValueError(f"learned closure: expected {_cl.NPARAM} parameters, got {flat.shape}")
        d = self.domain
        delta = p.delta if p.delta is not None else math.sqrt(
            d.size(1) / (2 * self.k1max + 2) * d.size(2) / (2 * self.k2max + 2))
        return flat, float(delta), wall_feature(d, self.n3, delta), int(bool(p.guard))

    def _arr(self, values):
        a = np.array(values, dtype=np.float64)
        self._keep.append(a)
        return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))

    def _config(self):
        """The constants of init_process for each process (src/processes/*.jl)."""
        d, n3 = self.domain, self.n3
        alpha = lambda z: n3 / d.Dvmap(z)  # derivatives.jl:10-19
        c = BldConfig()
        c.abi_version = _lib.BLD_ABI_VERSION
        c.precision = {"fp64": _lib.BLD_FP64, "fp32": _lib.BLD_FP32_MIXED}[self.precision]
        c.k1max, c.k2max, c.n3 = self.k1max, self.k2max, n3
        c.L1, c.L2 = d.hsize
        c.alpha_c = self._arr([alpha(z) for z in _zeta(n3, "C")])
        c.alpha_i = self._arr([alpha(z) for z in _zeta(n3, "C", neighbors=True)])
        for fi, f in enumerate(("vel1", "vel2", "vel3")):
            for side, (b, z) in enumerate(((d.lower_boundary, 0.0), (d.upper_boundary, 1.0))):
                kind, value = _bctype(b, f)
                bc = c.bc[fi][side]
                bc.kind = _lib.BLD_BC_DIRICHLET if kind == "dirichlet" else _lib.BLD_BC_NEUMANN
                bc.value = value
                bc.difference = value * (d.Dvmap(z) / n3) if kind == "neumann" else 0.0
        for p in self.processes:
            if isinstance(p, MomentumAdvection):
                c.advection = 1
                c.adv_n1, c.adv_n2 = pdsize((self.k1max, self.k2max), p.dealiasing)
            elif isinstance(p, MolecularDiffusion):
                fi = FIELD_INDEX[p.field]
                if c.diffusion[fi]:
                    raise NotImplementedError("more than one diffusion process per field")
                c.diffusion[fi] = 1
                c.diffusivity[fi] = float(p.diffusivity)
            elif isinstance(p, Pressure):
                c.pressure = 1
            elif isinstance(p, ConstantSource):
                c.source[FIELD_INDEX[p.field]] += float(p.strength)
            elif isinstance(p, ConstantMean):
                if NODES[p.field] != "C":
                    raise NotImplementedError("The constant-mean source currently does not handle I-nodes correctly")
                fi = FIELD_INDEX[p.field]
                c.mean[fi] = 1
                c.mean_value[fi] = float(p.mean_value)
                c.mean_weight = self._arr([1.0 / alpha(z) / d.size(3) for z in _zeta(n3, "C")])
            elif isinstance(p, StaticSmagorinskyModel):
                if c.sgs:
                    raise NotImplementedError("more than one SGS process")
                self._config_sgs(c, p)
            elif isinstance(p, LearnedSGSModel):
                if c.sgs:
                    raise NotImplementedError("more than one SGS process")
                self._config_sgs(c, p.fallback)
                self.closure = p
            else:
                raise NotImplementedError(f"process {type(p).__name__} is not part of the GPU step engine")
        return c

    def _config_sgs(self, c, p):
        """static_smagorinsky.jl:44-123."""
        d, n3 = self.domain, self.n3
        c.sgs = 1
        c.sgs_n1, c.sgs_n2 = pdsize((self.k1max, self.k2max), p.dealiasing)
        D1 = d.size(1) / (2 * self.k1max + 2)
        D2 = d.size(2) / (2 * self.k2max + 2)
        zc, zi = _zeta(n3, "C"), _zeta(n3, "I")
        lc = [p.model_constant * np.cbrt(D1 * D2 * (d.Dvmap(z) / n3)) for z in zc]
        li = [p.model_constant * np.cbrt(D1 * D2 * (d.Dvmap(z) / n3)) for z in zi]
        x3min, x3max = d.vmap(0.0), d.vmap(1.0)

        def mixing(x3, wall):  # :125-127
            if isinstance(wall, FreeSlipBoundary):
                return math.inf
            if isinstance(wall, RoughWall):
                return x3 * wall.von_karman_constant
            raise ValueError("SGS model only supports rough-wall and free-slip boundaries")

        if p.wall_damping:
            N = p.wall_damping_exponent
            ml = lambda x3: min(mixing(x3 - x3min, d.lower_boundary), mixing(x3max - x3, d.upper_boundary))
            lc = [(a ** (-N) + ml(d.vmap(z)) ** (-N)) ** (-1 / N) for a, z in zip(lc, zc)]
            li = [(a ** (-N) + ml(d.vmap(z)) ** (-N)) ** (-1 / N) for a, z in zip(li, zi)]
        c.sgs_lc = self._arr(lc)
        c.sgs_li = self._arr(li)
        dx3 = (d.vmap(zc[0]) - x3min, x3max - d.vmap(zc[-1]))
        for side, (b, sign) in enumerate(((d.lower_boundary, 1.0), (d.upper_boundary, -1.0))):
            if isinstance(b, RoughWall):
                lg = math.log(dx3[side] / b.roughness)
                c.sgs_wall[side] = _lib.BLD_WALL_ROUGH
                c.sgs_factor_strain[side] = sign * (1 / (2 * dx3[side] * lg))
                c.sgs_factor_stress[side] = sign * (-b.von_karman_constant ** 2 / lg ** 2)
            elif isinstance(b, FreeSlipBoundary):
                c.sgs_wall[side] = _lib.BLD_WALL_FREESLIP
            else:
                raise ValueError("SGS model only supports rough-wall and free-slip boundaries")


class Model:
    """A discretised model on one GPU (BoundaryLayerDynamics.jl:72-97).

    ``precision`` is ``"fp64"`` (plain form) or ``"fp32"`` (float32 fields with
    float64 pressure elimination, (0,0) mode and forcing)."""

    def __init__(self, resolution, domain, processes, precision="fp64"):
        disc = Discretization(resolution, domain, processes, precision)
        lib = _lib.load()
        h = ctypes.c_void_p()
        rc = lib.bld_create(ctypes.byref(disc.config), ctypes.byref(h))
        if rc != 0:
            raise BldError(rc, lib.bld_last_error(None).decode())
        self._attach(disc, lib, h, 0, owner=True)
        self._apply_closure()

    def _apply_closure(self):
        args = self.disc.closure_args()
        if args is not None:
            flat, delta, xw, flags = args
            self._use()
            self._check(self._lib.bld_sgs_closure(self._h, fptr(flat), delta, dptr(xw), flags))

    def _attach(self, disc, lib, h, index, owner):
        self.disc = disc
        self.k1max, self.k2max, self.n3 = disc.k1max, disc.k2max, disc.n3
        self.domain, self.processes, self.precision, self.fields = (disc.domain, disc.processes,
                                                                    disc.precision, disc.fields)
        self._lib, self._h, self._inst, self._owner = lib, h, index, owner

    @classmethod
    def _member(cls, disc, lib, h, index, ensemble):
        m = cls.__new__(cls)
        m._attach(disc, lib, h, index, owner=False)
        m._ensemble = ensemble  # keeps the engine alive
        return m

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and getattr(self, "_owner", False):
            self._lib.bld_destroy(h)
            self._h = None

    @property
    def fdsize(self):
        return (self.k1max + 1, 2 * self.k2max + 1)

    def nlayers(self, field):
        return self.n3 if NODES[field] == "C" else self.n3 - 1

    def _use(self):
        self._check(self._lib.bld_select_instance(self._h, self._inst))

    # -- engine calls ------------------------------------------------------
    def _check(self, rc):
        if rc != 0:
            raise BldError(rc, self._lib.bld_last_error(self._h).decode())

    def _empty(self, f):
        return np.zeros((*self.fdsize, self.nlayers(f)), dtype=np.complex128, order="F")

    def _triple(self, arrays):
        return [dptr(arrays.get(f)) for f in ("vel1", "vel2", "vel3")]

    @property
    def state(self):
        """Copy of the state (model.state of the reference) as host arrays."""
        out = {f: self._empty(f) for f in ("vel1", "vel2", "vel3")}
        self._use()
        self._check(self._lib.bld_get_state(self._h, *self._triple(out)))
        return {f: out[f] for f in self.fields}

    def set_state(self, **fields):
        arrays = {f: np.asfortranarray(np.asarray(v, dtype=np.complex128)) for f, v in fields.items()}
        for f, a in arrays.items():
            if a.shape != (*self.fdsize, self.nlayers(f)):
                raise ValueError(f"{f}: expected shape {(*self.fdsize, self.nlayers(f))}, got {a.shape}")
        self._use()
        self._check(self._lib.bld_set_state(self._h, *self._triple(arrays)))

    def rates(self):
        """compute_rates! of the current state (Processes.jl:39-62)."""
        out = {f: self._empty(f) for f in ("vel1", "vel2", "vel3")}
        self._use()
        self._check(self._lib.bld_rates(self._h, *self._triple(out)))
        return {f: out[f] for f in self.fields}

    def project(self):
        """apply_projections! (Processes.jl:67-75)."""
        self._check(self._lib.bld_project(self._h))

    def default_size(self):
        return pdsize((self.k1max, self.k2max), None)

    def coordinates(self, field, dim):
        n1, n2 = self.default_size()
        if dim == 1:
            return self.domain.hsize[0] * np.arange(n1) / n1
        if dim == 2:
            return self.domain.hsize[1] * np.arange(n2) / n2
        return np.array([self.domain.vmap(z) for z in _zeta(self.n3, NODES[field])])

    def set_physical(self, field, values, centered=False):
        """set_field! at the default physical size (physical_space.jl:253-269);
        ``centered``: the values sit at the cell centres in x1 and x2."""
        a = np.asfortranarray(np.asarray(values, dtype=np.float64))
        n1, n2 = self.default_size()
        if a.shape != (n1, n2, self.nlayers(field)):
            raise ValueError("physical values must have the default physical size")
        self._use()
        self._check(self._lib.bld_set_field_physical(self._h, FIELD_INDEX[field], dptr(a), int(bool(centered))))

    def initialize(self, **ics):
        """State.jl:11-31: evaluate functions (or constants) on the default grid."""
        for f in ics:
            if f not in self.fields:
                raise KeyError(f"Cannot initialize `{f}`: unknown field")
        n1, n2 = self.default_size()
        for f in ("vel1", "vel2", "vel3"):
            fn = ics.get(f, 0.0)
            X = np.meshgrid(self.coordinates(f, 1), self.coordinates(f, 2), self.coordinates(f, 3), indexing="ij")
            if np.isscalar(fn):
                vals = np.full(X[0].shape, float(fn))
            else:
                vals = _evaluate(fn, *X)
            self.set_physical(f, vals)

    def __getitem__(self, field):
        """Default-size physical values (State.jl:81-85)."""
        return self.get_physical(field)

    def get_physical(self, field, centered=False):
        """get_field at the default physical size (physical_space.jl:297-312)."""
        n1, n2 = self.default_size()
        out = np.zeros((n1, n2, self.nlayers(field)), dtype=np.float64, order="F")
        self._use()
        self._check(self._lib.bld_get_field_physical(self._h, FIELD_INDEX[field], dptr(out), int(bool(centered))))
        return out

    # -- mean profiles on the device (MeanProfiles, logging.jl:310-525) ---------
    def enable_profiles(self, names):
        """Collect the named profiles (see MeanProfiles) from now on; () stops."""
        mask = 0
        for n in names:
            if n not in _lib.PROFILES:
                raise ValueError(f"unknown profile {n!r}")
            mask |= 1 << _lib.PROFILES.index(n)
        self._use()
        self._check(self._lib.bld_profiles_enable(self._h, mask))
        self._profiles = tuple(names)

    def read_profiles(self, reset=False):
        """({name: time integral since the last reset}, steps integrated)."""
        out, steps = {}, ctypes.c_int64()
        self._use()
        for n in getattr(self, "_profiles", ()):
            k = _lib.PROFILES.index(n)
            a = np.zeros(self.n3 + 1 if n in _lib.I_PROFILES else self.n3)
            self._check(self._lib.bld_profiles_read(self._h, k, dptr(a), ctypes.byref(steps), 0))
            out[n] = a
        if reset and out:
            self._check(self._lib.bld_profiles_read(self._h, _lib.PROFILES.index(self._profiles[0]), None,
                                                    ctypes.byref(steps), 1))
        return out, steps.value

    # -- snapshots (CBD files, fileio.jl; Snapshots, logging.jl:526-602) ------
    def _xranges(self, field, centered):
        from . import cbd
        n1, n2 = self.default_size()
        d = self.domain
        x1 = cbd.hrange(d.hsize[0], n1, centered)
        x2 = cbd.hrange(d.hsize[1], n2, centered)
        x3 = np.array([d.vmap(z) for z in _zeta(self.n3, NODES[field])])
        xmin = (0.0, 0.0, d.vmap(0.0))
        xmax = (d.hsize[0], d.hsize[1], d.vmap(1.0))
        return x1, x2, x3, xmin, xmax

    def write_snapshot(self, path, centered=True, precision=None):
        """One `state-NNN` folder of the Snapshots output (logging.jl:584-602):
        a CBD file per field with the default-size physical values (cell
        centres with ``centered``), in the model precision unless given."""
        from . import cbd
        dtype = np.dtype(precision if precision is not None else
                         (np.float32 if self.precision == "fp32" else np.float64))
        for f in self.fields:
            x1, x2, x3, xmin, xmax = self._xranges(f, centered)
            cbd.write_cbd(dtype, os.path.join(path, f"{f}.cbd"), self.get_physical(f, centered),
                          x1, x2, x3, xmin, xmax)

    def initialize_from(self, path):
        """initialize!(model, path) (State.jl:33-66): read `<field>.cbd` of each
        field (missing files set the field to zero), check that the file fits
        the domain and grid, and set the state; the files must have the default
        physical size of this model."""
        from . import cbd
        n1, n2 = self.default_size()
        for f in self.fields:
            fn = os.path.join(path, f"{f}.cbd")
            if not os.path.isfile(fn):
                warnings.warn(f"No initial values found for `{f}`, set to zero.")
                self.set_physical(f, np.zeros((n1, n2, self.nlayers(f)), order="F"))
                continue
            xmin, xmax, x1, x2, x3, data = cbd.read_cbd(fn)
            _, _, x3ref, xmin_ref, xmax_ref = self._xranges(f, False)
            if not (np.allclose(xmin, xmin_ref) and np.allclose(xmax, xmax_ref)):
                raise ValueError(f"{fn}: domain extent does not match the model")
            if data.shape[:2] != (n1, n2):
                raise NotImplementedError(f"{fn}: snapshot size {data.shape[:2]} differs from the default "
                                          f"physical size {(n1, n2)}")
            centered = x1[0] != 0
            if not (np.allclose(x1, cbd.hrange(xmax[0] - xmin[0], n1, centered))
                    and np.allclose(x2, cbd.hrange(xmax[1] - xmin[1], n2, centered))):
                raise ValueError(f"{fn}: x1/x2 positions are not a uniform grid")
            if x3.shape != x3ref.shape or not np.allclose(x3, x3ref):
                raise ValueError(f"{fn}: x3 positions do not match the model grid")
            self.set_physical(f, data, centered)

    # -- time integration ----------------------------------------------------
    def init_problem(self, t1=0.0):
        self._check(self._lib.bld_init_problem(self._h, t1))

    def set_method(self, method):
        """ODE method of the step calls (an instance or class of SSPRK33,
        SSPRK22, Euler, AB2); applies to every model of an ensemble."""
        code = getattr(method, "code", None)
        if code is None:
            raise NotImplementedError(f"ODE method {method!r} is not part of the GPU step engine")
        self._check(self._lib.bld_set_method(self._h, code))

    def step(self, dt, nsteps, t=0.0, t_end=0.0, graph=True):
        """nsteps × (perform_step! + rate!); returns the new time.  A
        non-finite rate raises "The simulation has diverged." with the time the
        reference reaches before it raises as the error's ``t``."""
        tt = ctypes.c_double(t)
        fn = self._lib.bld_step_graph if graph else self._lib.bld_step
        rc = fn(self._h, dt, nsteps, ctypes.byref(tt), t_end)
        if rc != 0:
            err = BldError(rc, self._lib.bld_last_error(self._h).decode())
            err.t = tt.value
            raise err
        return tt.value

    def time_steps(self, dt, nsteps, graph=True):
        ms = ctypes.c_double()
        self._check(self._lib.bld_time_steps(self._h, dt, nsteps, 1 if graph else 0, ctypes.byref(ms)))
        return ms.value

    def autotune(self, reps=5):
        """FFT batch autotuner (bld_autotune); recomputes the rates of the state."""
        self._check(self._lib.bld_autotune(self._h, reps))

    TRANSFORM_PATHS = {"two-pass": 0, "whole-plane": 3, "half-plane": 7, "whole-plane-c2r": 1, "whole-plane-r2c": 2}

    def set_transform_path(self, path):
        """Force the transform path of every model space (bld_set_transform_path):
        "whole-plane" (c2r and r2c, 1024-thread blocks for the largest planes),
        "half-plane" (whole-plane, with half-plane c2r blocks where a plane needs
        1024 threads), "two-pass", or only one direction whole-plane."""
        self._check(self._lib.bld_set_transform_path(self._h, self.TRANSFORM_PATHS[path]))

    def tuning(self):
        """{pass: [(batch, µs) per model space]} chosen by the autotuner."""
        out = {}
        for k, name in enumerate(("c2r_x2", "c2r_x1", "r2c_x1", "r2c_x2")):
            out[name] = []
            for e in range(16):
                b, us = ctypes.c_int(), ctypes.c_double()
                if self._lib.bld_tuning(self._h, k, e, ctypes.byref(b), ctypes.byref(us)) != 0:
                    break
                out[name].append((b.value, round(us.value, 2)))
        b, us = ctypes.c_int(), ctypes.c_double()
        self._check(self._lib.bld_tuning(self._h, 4, 0, ctypes.byref(b), ctypes.byref(us)))
        out["stage_chunk"] = (b.value, round(us.value, 2))
        self._check(self._lib.bld_tuning(self._h, 5, 0, ctypes.byref(b), ctypes.byref(us)))
        out["transform_path"] = ({"c2r": ("half-plane" if b.value & 4 else "whole-plane") if b.value & 1 else "two-pass",
                                  "r2c": "whole-plane" if b.value & 2 else "two-pass"}, round(us.value, 2))
        return out

    def set_profiling(self, on):
        self._check(self._lib.bld_set_profiling(self._h, 1 if on else 0))

    def sample_closure(self, filter_width):
        """Training samples of the current state for the learned closure: box
        filter of `filter_width` cells in x1 and x2 on each C-layer of the
        default physical grid (Δ = filter_width × the geometric mean spacing).
        Returns (x, t, Δ): float32 [n, 16] features and targets (t[:, :6] the
        normalised deviatoric stress, t[:, 15] its scale Δ²|ḡ|²)."""
        d = self.domain
        delta = filter_width * math.sqrt(d.size(1) / (2 * self.k1max + 2) * d.size(2) / (2 * self.k2max + 2))
        xw = wall_feature(d, self.n3, delta)
        n = ctypes.c_int64()
        self._use()
        self._check(self._lib.bld_closure_sample(self._h, filter_width, delta, dptr(xw), None, None, 0,
                                                 ctypes.byref(n)))
        x = np.zeros((n.value, 16), dtype=np.float32)
        t = np.zeros((n.value, 16), dtype=np.float32)
        self._check(self._lib.bld_closure_sample(self._h, filter_width, delta, dptr(xw), fptr(x), fptr(t), n.value,
                                                 ctypes.byref(n)))
        return x, t, delta

    def time_kernel(self, name, reps=20):
        """(µs per dispatch, dispatches per right-hand side) of kernel class
        `name` replayed back to back (bld_time_kernel)."""
        us, n = ctypes.c_double(), ctypes.c_int64()
        self._check(self._lib.bld_time_kernel(self._h, _lib.KERNEL_KINDS.index(name), reps, ctypes.byref(us),
                                              ctypes.byref(n)))
        return us.value, n.value

    def kernel_times(self):
        out = {}
        for k, name in enumerate(_lib.KERNEL_KINDS):
            ms, n = ctypes.c_double(), ctypes.c_int64()
            self._check(self._lib.bld_kernel_times(self._h, k, ctypes.byref(ms), ctypes.byref(n)))
            out[name] = (ms.value, n.value)
        return out


def _evaluate(fn, X1, X2, X3):
    try:
        v = np.asarray(fn(X1, X2, X3), dtype=float)
        if v.shape == X1.shape:
            return v
    except (TypeError, ValueError):
        pass
    out = np.empty(X1.shape)
    for idx in np.ndindex(X1.shape):
        out[idx] = fn(X1[idx], X2[idx], X3[idx])
    return out


def _approxdiv(a, b):  # util.jl:19-24
    r = a / b
    n = int(round(r))
    if not (r == n or abs(r - n) <= math.sqrt(np.finfo(float).eps) * max(abs(r), abs(n))):
        raise ValueError("Division has a non-integer result")
    return n


class MeanProfiles:
    """MeanProfiles output (logging.jl:310-525): time averages (trapezoidal
    rule over the samples after each step) of horizontal means, collected on
    the GPU.  ``profiles`` from vel1-3, adv1-3, sgs11, sgs12, sgs13, sgs22,
    sgs23, sgs33; at every multiple of ``output_frequency`` (and at the end)
    the averages since the previous output are written to `path`-NN.npz, one
    array per profile plus `intervals` and `timespan` (the reference writes
    the same datasets and attributes to HDF5, which is not available here)."""

    def __init__(self, profiles=("vel1", "vel2", "vel3"), path=os.path.join("output", "profiles"),
                 output_frequency=None):
        self.profiles, self.path, self.output_frequency = tuple(profiles), path, output_frequency

    def due(self, t):
        if self.output_frequency is None or t == 0:
            return False
        n = round(t / self.output_frequency)
        return math.isclose(n, t / self.output_frequency, rel_tol=1e-9, abs_tol=1e-12)

    def start(self, model, t):
        model.enable_profiles(self.profiles)
        self.t0 = t

    def flush(self, model, t):
        integrals, steps = model.read_profiles(reset=True)
        if steps == 0 or not t > self.t0:
            return None
        means = {k: v / (t - self.t0) for k, v in integrals.items() if k in self.profiles}
        d = os.path.dirname(self.path)
        if d:
            os.makedirs(d, exist_ok=True)
        i = 1
        while os.path.exists(f"{self.path}-{i:02d}.npz"):
            i += 1
        fn = f"{self.path}-{i:02d}.npz"
        np.savez(fn, intervals=np.int64(steps), timespan=np.array([self.t0, t]), **means)
        self.t0 = t
        return fn


def _layer_physical(fd, n1, n2):
    """Physical values of one layer of spectral coefficients (k1 0..K1-1,
    k2 in the order 0..k2max, -k2max..-1) on an n1 x n2 grid: zero-padded,
    unnormalised inverse transform (physical_space.jl:277-312)."""
    K1, K2 = fd.shape
    k1m, k2m = min(n1 // 2 + 1, K1) - 1, (min(n2, K2) - 1) // 2
    buf = np.zeros((n1 // 2 + 1, n2), dtype=np.complex128)
    buf[: k1m + 1, : k2m + 1] = fd[: k1m + 1, : k2m + 1]
    if k2m > 0:
        buf[: k1m + 1, n2 - k2m:] = fd[: k1m + 1, K2 - k2m:]
    return np.fft.irfftn(buf, s=(n2, n1), axes=(1, 0), norm="forward")


class StepTimer:
    """StepTimer output (logging.jl:76-127): the wall time and the simulation
    time after every `frequency`-th step (and of the initial state), written
    as JSON to `path` at the end: {"frequency", "wallTime" (seconds since the
    timer was made), "simulationTime"}.  The engine steps in chunks of
    `frequency` steps while a StepTimer is active, so each stamp is taken
    after its step has finished on the GPU."""

    def __init__(self, path=os.path.join("output", "timestamps.json"), frequency=1):
        import time
        self.path, self.frequency = path, int(frequency)
        self.init_time = time.time()
        self.wall, self.sim = [], []

    def stamp(self, t):
        import time
        self.wall.append(time.time() - self.init_time)
        self.sim.append(float(t))

    def flush(self):
        d = os.path.dirname(self.path)
        if d:
            os.makedirs(d, exist_ok=True)
        with open(self.path, "w") as io:
            io.write("{\n")
            io.write(f'  "frequency": {self.frequency},\n')
            io.write('  "wallTime": [' + ", ".join(repr(v) for v in self.wall) + "],\n")
            io.write('  "simulationTime": [' + ", ".join(repr(v) for v in self.sim) + "]\n")
            io.write("}\n")


class ProgressMonitor:
    """ProgressMonitor output (logging.jl:129-308): at most every `frequency`
    seconds of wall time (checked after each chunk of steps, and at the end) a
    report with the progress bar and ETA, the simulation time, the mean
    velocity (the z-weighted mean of the (0,0) modes), the mean rough-wall
    stresses where the model has an SGS process, the mean and turbulent
    kinetic energy (horizontal means of the physical fields, summed over the
    layers) and the advective Courant number Δt/min(Δx3/max|u3|) (with
    `tstep`) or the advective time scale; a dot after every chunk in between.
    Physical values are taken at the default grid (the reference uses the
    largest physical grid of the model).  `out` receives the text."""

    def __init__(self, frequency=10, description="Progress:", tstep=None, out=None, chunk=64):
        import sys
        self.frequency, self.description, self.tstep = frequency, description, tstep
        self.out = sys.stdout if out is None else out
        self.chunk = int(chunk)
        self.timestamp = math.nan
        self.progress = math.nan
        self.remaining = math.nan
        self.reports = []

    def start(self, t1, t2):
        self.range = (t1, t2)

    def samples(self, model, t):
        d, n3 = model.domain, model.n3
        dz = lambda z: d.Dvmap(z) / n3
        cw = np.array([dz(z) for z in _zeta(n3, "C")]) / d.size(3)
        iw = np.array([dz(z) for z in _zeta(n3, "I")]) / d.size(3)
        st = model.state
        out = [("Simulation Time", (t,))]
        means = []
        for f, w in (("vel1", cw), ("vel2", cw), ("vel3", iw)):
            means.append(float(np.sum(st[f][0, 0, :].real * w)) if f in st else 0.0)
        out.append(("Mean Velocity", tuple(means)))
        sgs = [p for p in model.processes if isinstance(p, (StaticSmagorinskyModel, LearnedSGSModel))]
        if sgs:
            # rough-wall stress at the first C-layer next to each wall, on the
            # dealiased grid of the SGS terms (static_smagorinsky.jl:176-184)
            c = model.disc.config
            vals = {0: [], 1: []}
            for side, lev in ((0, 0), (1, n3 - 1)):
                if c.sgs_wall[side] == _lib.BLD_WALL_ROUGH:
                    a = _layer_physical(st["vel1"][:, :, lev], c.sgs_n1, c.sgs_n2)
                    b = _layer_physical(st["vel2"][:, :, lev], c.sgs_n1, c.sgs_n2)
                    mag = np.sqrt(a * a + b * b)
                    vals[0].append(float(np.mean(c.sgs_factor_stress[side] * a * mag)))
                    vals[1].append(float(np.mean(c.sgs_factor_stress[side] * b * mag)))
                else:
                    vals[0].append(0.0)
                    vals[1].append(0.0)
            out.append(("Wall Stress τ₁₃", tuple(vals[0])))
            out.append(("Wall Stress τ₂₃", tuple(vals[1])))
        u, v = model.get_physical("vel1"), model.get_physical("vel2")
        u1 = u.mean(axis=(0, 1))
        u2 = v.mean(axis=(0, 1))
        mke = u1 ** 2 + u2 ** 2
        ke = (u ** 2).mean(axis=(0, 1)) + (v ** 2).mean(axis=(0, 1))
        out.append(("Mean KE", (float(np.sum(mke)),)))
        out.append(("Turbulent KE", (float(np.sum(ke - mke)),)))
        w = model.get_physical("vel3")
        dx3i = np.array([dz(z) for z in _zeta(n3, "I")])
        wmax = np.abs(w).max(axis=(0, 1))
        with np.errstate(divide="ignore"):
            dtmin = float(np.min(dx3i / wmax)) if wmax.size else math.inf
        if self.tstep is not None:
            out.append(("Adv. Courant Number", (self.tstep / dtmin,)))
        else:
            out.append(("Advective Timescale", (dtmin,)))
        return out

    def check(self, model, t):
        """prepare_samples!, log_state! and process_log! (logging.jl:188-306)
        at time t after a chunk of steps: a report when due, else a dot."""
        import time
        progress = (t - self.range[0]) / (self.range[1] - self.range[0]) if self.range[1] > self.range[0] else 1.0
        now = time.time()
        due = math.isnan(self.timestamp) or now - self.timestamp >= self.frequency or progress >= 1
        if not due:
            self.out.write("·")
            self.out.flush()
            return None
        self.out.write("\n")
        rate = (progress - self.progress) / (now - self.timestamp) if not math.isnan(self.timestamp) else math.nan
        self.remaining = (1 - progress) / rate if rate and rate > 0 else math.nan
        self.timestamp, self.progress = now, progress
        samples = self.samples(model, t)
        eta = "unknown"
        if self.remaining < 9e9:
            sec = int(math.ceil(self.remaining))
            eta = f"{sec // 3600}:{(sec // 60) % 60:02d}:{sec % 60:02d}"
        L, blocks = 25, " ▏▎▍▌▋▊▉█"
        ip = int(round(progress * L * 8))
        bar = "".join(blocks[min(max(1 + ip - i, 1), 9) - 1] for i in range(0, 8 * L, 8))
        msg = f"{self.description} {100 * progress:3.0f}%▕{bar}▏ ETA: {eta}"
        wt = len(msg)
        wl = max(len(k) for k, _ in samples) + 2
        lines = ["─" * wt, msg, "─" * wl + "┬" + "─" * (wt - wl - 1)]
        for label, vals in samples:
            lines.append(label.rjust(wl - 1) + " ╎ " + ", ".join(f"{v:.3g}" for v in vals))
        lines.append("─" * wl + "┴" + "─" * (wt - wl - 1))
        self.out.write("\n".join(lines) + "\n")
        self.out.flush()
        self.reports.append(dict(samples))
        return samples


class Snapshots:
    """Snapshots output (logging.jl:526-602): the state as CBD files in
    `path`/state-NNN/ at every multiple NNN of `frequency` (the initial state is
    not written), values at the cell centres unless ``centered=False``."""

    def __init__(self, frequency, path=os.path.join("output", "snapshots"), centered=True, precision=None):
        self.frequency, self.path, self.centered, self.precision = frequency, path, centered, precision

    def due(self, t):
        if t == 0:
            return None
        n = round(t / self.frequency)
        return n if math.isclose(n, t / self.frequency, rel_tol=1e-9, abs_tol=1e-12) else None

    def write(self, model, n):
        model.write_snapshot(os.path.join(self.path, f"state-{n:03d}"), self.centered, self.precision)


def evolve(model, tspan, dt, method=None, graph=True, output=()):
    """evolve! (BoundaryLayerDynamics.jl:190-219).

    The state is projected and its rate evaluated at t1 (ODEProblem,
    ODEMethods.jl:179-185), then nt steps of `method` (default SSPRK33) run on
    the GPU, each followed by the rate evaluation of ODEMethods.jl:231.
    ``output``: Snapshots objects, written after the steps that reach their
    times.  Returns the final time."""
    model.set_method(SSPRK33() if method is None else method)
    t1, t2 = (0.0, float(tspan)) if np.isscalar(tspan) else (float(tspan[0]), float(tspan[-1]))
    if not t1 <= t2:
        raise ValueError("Integration start time is larger than end time")
    outputs = list(output) if isinstance(output, (list, tuple)) else [output]
    for o in outputs:
        if not isinstance(o, (Snapshots, MeanProfiles, StepTimer, ProgressMonitor)):
            raise NotImplementedError(f"output {type(o).__name__} is not supported by the GPU engine")
    if sum(isinstance(o, MeanProfiles) for o in outputs) > 1:
        raise NotImplementedError("one MeanProfiles output per model")
    model.init_problem(t1)
    for o in outputs:
        if isinstance(o, MeanProfiles):
            o.start(model, t1)
        elif isinstance(o, StepTimer):
            o.stamp(t1)
        elif isinstance(o, ProgressMonitor):
            o.start(t1, t2)
            o.check(model, t1)
    if t2 <= t1:
        return t1
    nt = _approxdiv(t2 - t1, dt)
    if not outputs:
        return model.step(dt, nt, t=t1, t_end=t2, graph=graph)
    timers = [o for o in outputs if isinstance(o, StepTimer)]
    monitors = [o for o in outputs if isinstance(o, ProgressMonitor)]
    timed = [o for o in outputs if isinstance(o, (Snapshots, MeanProfiles))]
    # step to each output time (and each StepTimer stamp or ProgressMonitor
    # check) in one engine call
    t, done = t1, 0
    while done < nt:
        k = 1
        limit = min([nt - done] + [o.frequency - done % o.frequency for o in timers] + [m.chunk for m in monitors])
        while k < limit and not any(o.due(t1 + (done + k) * dt) for o in timed):
            k += 1
        t = model.step(dt, k, t=t, t_end=t2, graph=graph)
        done += k
        for o in timed:
            if isinstance(o, Snapshots):
                n = o.due(t)
                if n is not None:
                    o.write(model, n)
            elif o.due(t):
                o.flush(model, t)
        for o in timers:
            if done % o.frequency == 0:
                o.stamp(t)
        for o in monitors:
            o.check(model, t)
    for o in outputs:  # flush! at the end of the run (logging.jl:453-491)
        if isinstance(o, MeanProfiles):
            o.flush(model, t)
            model.enable_profiles(())
        elif isinstance(o, StepTimer):
            o.flush()
    return t


class Ensemble:
    """Independent models (uneven grids allowed) advanced together: one launch
    per kernel class covers every member and a whole SSPRK33 step of the
    ensemble is one HIP graph.  ``members[i]`` is a `Model` view of member i."""

    def __init__(self, specs, precision="fp32"):
        self.discs = [Discretization(res, dom, procs, precision) for res, dom, procs in specs]
        lib = _lib.load()
        cfgs = (BldConfig * len(self.discs))(*[d.config for d in self.discs])
        h = ctypes.c_void_p()
        rc = lib.bld_ensemble_create(len(self.discs), cfgs, ctypes.byref(h))
        if rc != 0:
            raise BldError(rc, lib.bld_last_error(None).decode())
        self._lib, self._h = lib, h
        self.members = [Model._member(d, lib, h, i, self) for i, d in enumerate(self.discs)]
        for m in self.members:
            m._apply_closure()

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._lib.bld_destroy(h)
            self._h = None

    def __len__(self):
        return len(self.members)

    def _check(self, rc):
        if rc != 0:
            raise BldError(rc, self._lib.bld_last_error(self._h).decode())

    def init_problem(self, t1=0.0):
        self._check(self._lib.bld_init_problem(self._h, t1))

    def step(self, dt, nsteps, graph=True):
        t = ctypes.c_double(0.0)
        fn = self._lib.bld_step_graph if graph else self._lib.bld_step
        self._check(fn(self._h, dt, nsteps, ctypes.byref(t), 0.0))

    set_method = Model.set_method

    def time_steps(self, dt, nsteps, graph=True):
        ms = ctypes.c_double()
        self._check(self._lib.bld_time_steps(self._h, dt, nsteps, 1 if graph else 0, ctypes.byref(ms)))
        return ms.value

    def autotune(self, reps=5):
        self._check(self._lib.bld_autotune(self._h, reps))

    def set_profiling(self, on):
        self._check(self._lib.bld_set_profiling(self._h, 1 if on else 0))

    kernel_times = Model.kernel_times

    def grid_points(self):
        return sum(d.n3 * (2 * d.k1max + 2) * (2 * d.k2max + 2) for d in self.discs)
