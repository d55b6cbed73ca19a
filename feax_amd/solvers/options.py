"""Solver option dataclasses (the feax.solvers.options API).

Field names and defaults follow feax/solvers/options.py:444-512 and
783-889.  Added fields (all defaulted, so feax call sites work unchanged):

KrylovSolverOptions
  precision  "fp64" (default; assembly, SpMV and CG in FP64 like feax's
             FEAX_X64 default); "mixed" (FP32 assembly and FP32 matrix in
             difference form, FP64 CG vectors and dot products, inside an
             FP64 defect correction whose residual b - K u is the FP64
             matrix-free residual kernel); "mixed-fp32vec" (same with FP32
             CG vectors).
  spmv       "auto" | "csr" | "block3": SpMV kernel for the CG iterations.
  check_every  iterations between host reads of the convergence flag.
  inner_tol, max_outer  mixed-precision inner CG tolerance and outer cap.
  inner_safety  mixed precision: a later inner solve stops at relative
             tolerance max(inner_tol, inner_safety * target / ||r||), the
             reduction the outer target still needs with a margin (capped
             at 0.5; 0 keeps inner_tol throughout).
  recycle    mixed precision: number of search directions of the first
             inner solve kept (FP32, recycle x n in HBM) to start the later
             inner solves from the Galerkin projection of their residual
             (0 disables).
  operator   "csr" (default: assembled BC-eliminated CSR, Jacobi-PCG) or
             "matfree" (feax's own default Krylov operator, the matrix-free
             J_bc v, with no preconditioner as in feax, whose Jacobi flag is
             ignored for a callable operator, feax/solvers/common.py:221-227).
             precision "fp64" only.
"""

from dataclasses import dataclass
from enum import Enum
from typing import Callable, Optional
import logging

logger = logging.getLogger(__name__)


class MatrixProperty(Enum):
    SPD = "spd"
    SYMMETRIC = "symmetric"
    GENERAL = "general"


@dataclass(frozen=True)
class AbstractSolverOptions:
    verbose: bool = False
    check_convergence: bool = False
    convergence_threshold: float = 0.1

    def uses_x0(self) -> bool:
        raise NotImplementedError


@dataclass(frozen=True)
class NewtonOptions:
    tol: float = 1e-6
    rel_tol: float = 1e-8
    max_iter: int = 100
    line_search_max_backtracks: int = 30
    line_search_c1: float = 1e-4
    line_search_rho: float = 0.5
    raise_on_line_search_failure: bool = True
    verbose: bool = False


@dataclass(frozen=True)
class KrylovSolverOptions(AbstractSolverOptions):
    solver: str = "auto"
    tol: float = 1e-10
    atol: float = 1e-10
    maxiter: int = 10000
    preconditioner: Optional[Callable] = None
    use_jacobi_preconditioner: bool = False
    jacobi_shift: float = 1e-12
    restart: Optional[int] = None
    x0_fn: Optional[Callable] = None
    precision: str = "fp64"
    spmv: str = "auto"
    check_every: int = 32
    inner_tol: float = 1e-5
    inner_safety: float = 0.3
    max_outer: int = 30
    recycle: int = 3000
    operator: str = "csr"

    def __post_init__(self):
        valid = ("auto", "cg", "bicgstab", "gmres")
        if self.solver not in valid:
            raise ValueError(f"Invalid iterative solver: {self.solver}. Choose from {valid}")
        if self.precision not in ("fp64", "mixed", "mixed-fp32vec"):
            raise ValueError(f"Invalid precision: {self.precision}. Choose from ('fp64', 'mixed', 'mixed-fp32vec')")
        if self.spmv not in ("auto", "csr", "block3"):
            raise ValueError(f"Invalid spmv: {self.spmv}. Choose from ('auto', 'csr', 'block3')")
        if self.operator not in ("csr", "matfree"):
            raise ValueError(f"Invalid operator: {self.operator}. Choose from ('csr', 'matfree')")

    def uses_x0(self) -> bool:
        return True


@dataclass(frozen=True)
class DirectSolverOptions(AbstractSolverOptions):
    """feax.solvers.options.DirectSolverOptions (options.py:595-635).  Every
    solver name runs the device banded Cholesky of the SPD J_bc
    (symmetric elimination): "auto" and "cudss" (feax's GPU solver), and the
    host solvers "spsolve", "umfpack", "cholmod" alike.  cudss_options and
    sksparse_options are accepted and unused; the forward factor is always
    reused by the adjoint solve (J_bc is symmetric), which feax does under
    reuse_factorization."""
    solver: str = "auto"
    cudss_options: Optional[object] = None
    sksparse_options: Optional[object] = None
    reuse_factorization: bool = False

    def __post_init__(self):
        valid = ("auto", "cudss", "spsolve", "umfpack", "cholmod")
        if self.solver not in valid:
            raise ValueError(f"Invalid direct solver: {self.solver}. Choose from {valid}")

    def uses_x0(self) -> bool:
        return False


def resolve_iterative_solver(options, matrix_property):
    """'auto' -> cg / bicgstab / gmres by matrix property (options.py:837-889)."""
    if options.solver != "auto":
        return options
    name = {MatrixProperty.SPD: "cg", MatrixProperty.SYMMETRIC: "bicgstab",
            MatrixProperty.GENERAL: "gmres"}[matrix_property]
    import dataclasses
    logger.info(f"IterativeSolver auto: matrix_property={matrix_property.name} -> {name}")
    return dataclasses.replace(options, solver=name)


def detect_matrix_property(A, sym_tol: float = 1e-8, matrix_view=None) -> MatrixProperty:
    """SPD / SYMMETRIC / GENERAL of an assembled device CSRMatrix
    (feax/solvers/options.py:188-259).

    Symmetry: ||A x - A^T x|| / (||A x|| + 1e-30) < sym_tol for a seeded
    standard-normal x (feax draws x from jax.random.PRNGKey(42); the
    decision does not depend on the draw for an exactly symmetric or a
    clearly non-symmetric matrix), skipped for an UPPER or LOWER view.
    Definiteness (heuristic, as in feax): every diagonal entry positive.
    """
    import numpy as np
    import torch
    from .common import device_norm
    name = getattr(matrix_view, "name", matrix_view)
    if name in ("UPPER", "LOWER"):
        symmetric = True
    else:
        n = A.shape[1]
        x = torch.as_tensor(np.random.default_rng(42).standard_normal(n), dtype=A.data.dtype,
                            device=A.data.device)
        Ax = A @ x
        ATx = A.rmatvec(x)
        symmetric = device_norm(Ax - ATx) / (device_norm(Ax) + 1e-30) < sym_tol
    if not symmetric:
        return MatrixProperty.GENERAL
    if bool((A.diagonal() > 0).all()):
        return MatrixProperty.SPD
    return MatrixProperty.SYMMETRIC


def resolve_direct_solver(options, matrix_property, matrix_view=None):
    """'auto' -> the device solver (feax/solvers/options.py:637-780 picks
    cuDSS on a GPU backend).  The device banded factorisation is Cholesky /
    L D L^T of a symmetric matrix, so a GENERAL matrix is refused."""
    if matrix_property == MatrixProperty.GENERAL:
        raise NotImplementedError("the direct solver factors symmetric matrices; the detected matrix is "
                                  "GENERAL (use KrylovSolverOptions(solver='gmres' or 'bicgstab'))")
    if options.solver != "auto":
        return options
    import dataclasses
    logger.info(f"DirectSolver auto: matrix_property={matrix_property.name} -> cudss")
    return dataclasses.replace(options, solver="cudss")
