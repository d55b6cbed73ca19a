"""feax_amd: MI355X-native HEX8 assembly + Krylov path with the feax API.

Drop-in for the KrylovSolverOptions route of feax.create_solver: Problem,
DirichletBCSpec / DirichletBCConfig, TracedParams, TracedStructure,
KrylovSolverOptions, DirectSolverOptions, create_solver, Solution,
zero_like_initial_guess, CSRMatrix.  All compute runs in libfeaxhip.so (hand-written HIP for
gfx950); the package itself holds only host-side structure and dispatch.
"""

from . import gene, materials
from .assembler import (TracedStructure, create_energy_fn, create_J_bc_csr_function, create_J_bc_csr_parametric,
                        create_res_bc_function, create_res_bc_parametric, create_res_J_bc_csr_parametric,
                        get_jacobian, get_jacobian_info, get_res)
from .csr import CSRMatrix
from .DCboundary import DirichletBC, DirichletBCConfig, DirichletBCSpec, dirichlet_bc_config
from .mesh import Mesh, StructuredGrid, hex20_mesh
from .problem import MatrixView, Problem
from .solution import Solution
from .solver import create_solver
from .solvers.gmg import GeometricMultigrid
from .solvers.options import DirectSolverOptions, KrylovSolverOptions, MatrixProperty, NewtonOptions
from .traced_params import TracedParams
from .utils import zero_like_initial_guess

__version__ = "0.1.0"
from .solvers.common import (check_convergence, create_direct_solve_fn, create_iterative_solve_fn,
                             create_jacobi_preconditioner, create_linear_solve_fn, create_x0)
from .solvers.linear import create_linear_solver, linear_solve
from .solvers.options import (AbstractSolverOptions, detect_matrix_property, resolve_direct_solver,
                              resolve_iterative_solver)
