"""MI355X time-step engine for BoundaryLayerDynamics.jl (host package).

The compute path is ``libbld.so`` (HIP kernels for gfx950 behind the C ABI of
``include/bld.h``); this package mirrors the reference's Julia model API on
top of it.  Build the library with ``build.build()``.
"""

from .model import (  # noqa: F401
    Domain, SmoothWall, RoughWall, FreeSlipBoundary, CustomBoundary, SinusoidalMapping,
    MomentumAdvection, MolecularDiffusion, Pressure, ConstantSource, ConstantMean, StaticSmagorinskyModel,
    LearnedSGSModel, SSPRK33, SSPRK22, Euler, AB2,
    incompressible_flow, closedchannelflow, openchannelflow, Model, Ensemble, evolve, pdsize, Snapshots,
    MeanProfiles, StepTimer, ProgressMonitor,
)
from . import _lib  # noqa: F401
from .ldlt import BatchLDLt  # noqa: F401
from . import workloads  # noqa: F401
from . import closure  # noqa: F401
from . import cbd  # noqa: F401
