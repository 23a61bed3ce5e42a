"""heatflow_amd -- MI355X (gfx950) implementation of HeatFlow's per-timestep hot path.

H-matrix assembly, the Fourier/Cattaneo right-hand side and the CG solve of
ExeQuantCode/HeatFlow (src/heatflow/mod_evolve.f90:45-177) as hand-written HIP kernels behind
the C-ABI in include/heatflow_hip.h.  Python entry points:

  Simulation           the reference procedures (sparse_Hmatrix, initial_evolve, simulate,
                       linbcg) on one device context
  System/Params/Cells  the `inputs` state the path reads; read_inputs() reads param.in,
                       mat.in and system.in
  workloads            the synthetic benchmark grids
  autograd             torch.autograd.Function over the hand-written adjoint
"""
from ._lib import HeatFlowError, load as load_library  # noqa: F401
from .evolution import Simulation  # noqa: F401
from .system import (Cells, Geometry, Params, System, read_inputs, read_mat,  # noqa: F401
                     read_params, read_system)
