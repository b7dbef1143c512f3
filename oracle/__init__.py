"""FP64 CPU oracle for the Helmholtz solve-and-gradient path.

TEST INFRASTRUCTURE ONLY.  This package is the checker: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it, and only to compare against or to time a CPU baseline.  The product
path (``waveforminversionust_amd``) never imports it and fails loudly when its
HIP library is missing.

Everything here is a NumPy/SciPy restatement of the reference's MATLAB
algorithm, written from its text (each function cites the file:line it
follows, relative to the reference repository root):

* ``Runtime/solver/stencilOptParams.m``         -> :func:`helmholtz_fp64.stencil_opt_params`
* ``Runtime/solver/HelmholtzSolver.m:53-101``   -> :func:`helmholtz_fp64.pml_terms`
* ``Runtime/solver/ninePointCoefficients.m``    -> :func:`helmholtz_fp64.nine_point_coefficients`
* ``Runtime/solver/kwaveLdr9WavenumberSquared.m`` -> :func:`helmholtz_fp64.kwave_ldr9_wavenumber_squared`
* ``Runtime/solver/HelmholtzSolver.m:114-220``  -> :class:`helmholtz_fp64.HelmholtzFP64`
* ``Runtime/matlab/wust_oracle.m``              -> :func:`helmholtz_fp64.wust_oracle`
* ``Runtime/matlab/directional_curvature.m``    -> :func:`helmholtz_fp64.directional_curvature`
* ``Runtime/matlab/wust_geometry.m``            -> :func:`helmholtz_fp64.geometry`

How each function is pinned (the tests that hold it):

* stencil_opt_params: independent normal-equation solve and a dispersion
  bound on the fitted stencil (tests/test_oracle.py).
* pml_terms / nine_point_coefficients / assembly: a dense brute-force
  assembly written node by node from the Chen/Cheng/Feng/Wu finite-element
  form on 7x7 and 9x11 grids, and the homogeneous no-PML case, where the
  operator must be symmetric and reduce to the textbook 9-point stencil.
* solves: residual ||Hu-b||/||b|| <= 1e-12 and agreement with a dense LAPACK
  solve on small grids; the adjoint identity <Hx,y> = <x,H^H y>.
* mass_apply: <Mx,y> = <x,M^T y> with the border semantics.
* wust_oracle / directional_curvature: central finite differences of the
  loss (gradient) and of the linearised residual (curvature), batch
  invariance and loss-only equality, as in Runtime/tests/test_runtime.m.
* kwave_ldr9_wavenumber_squared: finite difference of q_eff in slowness.

The reference ships no golden vectors (its fixtures are regenerated by
running MATLAB, which is absent), so every pin above is a mathematical one.
"""
