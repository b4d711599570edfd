"""CPU oracle for the BoundaryLayerDynamics.jl SSPRK33 time step.

THIS PACKAGE IS TEST INFRASTRUCTURE, NOT PRODUCT CODE.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it, and only as the checker / the timed CPU restatement.  The GPU
engine under ``boundarylayerdynamics.jl_amd/`` never imports or calls it.

It is a float64 NumPy restatement of the reference Julia package
(efpl-columbia/BoundaryLayerDynamics.jl v1.0.0, ``src/``), written from the
reference sources read as text.  Julia is not available in this environment,
so the reference itself cannot be run; every function cites the reference
``file:line`` it restates.  Parity of each part is pinned by the reference's
own closed-form tests (re-expressed in ``tests/test_oracle_*.py``):

* grid layout ........ ``test/grid_test.jl``
* transforms ......... ``test/transform_test.jl`` + brute-force DFT
* vorticity .......... ``test/derivatives_test.jl``
* diffusion .......... ``test/diffusion_test.jl``
* advection .......... ``test/advection_test.jl`` (exactness, KE conservation)
* pressure ........... ``test/pressure_test.jl`` (LDLt vs dense, projection, KE)
* Smagorinsky ........ ``test/les_test.jl`` (wall stress, eddy viscosity)
* time integration ... ``test/ode_test.jl``, ``test/laminar_test.jl``

The FFT library of the reference (FFTW 3.3.10 through FFTW.jl 1.4.6, see
``benchmark/Manifest.toml:85-95``) is replaced by ``scipy.fft`` (pocketfft);
both compute the same unnormalised DFT sums, pinned by the closed-form
transform tests and a brute-force DFT check.
"""

from .model import (  # noqa: F401
    Domain, SmoothWall, RoughWall, FreeSlipBoundary, CustomBoundary, SinusoidalMapping,
    MomentumAdvection, MolecularDiffusion, Pressure, ConstantSource, ConstantMean,
    StaticSmagorinskyModel, incompressible_flow, Model, Grid, pdsize, wavenumbers,
    dx3factors, vrange, get_field, add_field, set_field, evolve, compute_rates,
    apply_projections, BatchLDLt,
)
from .ode import ODEProblem, solve, Euler, AB2, SSPRK22, SSPRK33, approxdiv  # noqa: F401
