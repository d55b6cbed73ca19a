"""The Solution return value on the GPU, mirroring the reference's
tests/test_solution.py: wraps and matches the flat solve, field / node_var
access, reuse as the next initial guess, gradients through .flat and
.field(0), and a thermal Solution passed straight into a thermo-elastic
solve's TracedParams."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _setup(dims=(4, 3, 3)):
    import feax_amd as fe
    mesh = fe.StructuredGrid(dims).to_mesh()

    class Elast(fe.Problem):
        def get_tensor_map(self):
            return fe.materials.IsotropicLinearElastic(nu=0.3)

    problem = Elast(mesh, vec=3, dim=3, ele_type="HEX8")
    E = np.full(problem.num_cells, 100.0)
    bc = fe.DirichletBCConfig([fe.DirichletBCSpec(lambda p: np.isclose(p[0], 0.0), "all", 0.0),
                               fe.DirichletBCSpec(lambda p: np.isclose(p[0], float(dims[0])), "x", 0.05)]
                              ).create_bc(problem)
    return mesh, problem, E, bc


def test_solution_wraps_reuses_and_differentiates():
    import torch
    import feax_amd as fe
    mesh, problem, E, bc = _setup()
    opts = fe.KrylovSolverOptions(solver="cg", tol=1e-12, atol=0.0, use_jacobi_preconditioner=True)
    ig = fe.zero_like_initial_guess(problem, bc)
    tp = fe.TracedParams(volume_vars=(E,))
    print("solution wrapper", flush=True)
    flat = fe.create_solver(problem, bc, opts, linear=True, return_solution=False)(tp, ig)
    solver = fe.create_solver(problem, bc, opts, linear=True)
    sol = solver(tp, ig)
    assert isinstance(sol, fe.Solution) and sol.layout == ((problem.num_total_nodes, 3),)
    assert torch.allclose(sol.flat, flat)
    assert torch.allclose(sol.field(0), flat.reshape(-1, 3)) and float(sol[3]) == float(flat[3])
    with pytest.raises(ValueError):
        sol.node_var()
    assert torch.allclose(fe.TracedParams.node_var_from_solution(problem, sol, component=0), flat.reshape(-1, 3)[:, 0])
    sol2 = solver(tp, sol)                       # a Solution as the next initial guess
    assert isinstance(sol2, fe.Solution) and float((sol2.flat - sol.flat).abs().max()) <= 1e-10
    Et = torch.tensor(E, device="cuda", requires_grad=True)
    (solver(fe.TracedParams(volume_vars=(Et,)), ig).flat ** 2).sum().backward()
    g1 = Et.grad.clone()
    Et.grad = None
    (solver(fe.TracedParams(volume_vars=(Et,)), ig).field(0) ** 2).sum().backward()
    assert torch.allclose(g1, Et.grad)


def test_thermal_solution_as_direct_coefficient():
    """The reference's test_solution_as_direct_coefficient_scalar: the thermal
    Solution itself goes into the thermo-elastic TracedParams; the heated
    bar expands in +x."""
    import feax_amd as fe
    mesh = fe.StructuredGrid((4, 3, 3)).to_mesh()

    class Thermal(fe.Problem):
        def get_tensor_map(self):
            return fe.materials.IsotropicConduction()

    class ThermoElastic(fe.Problem):
        def get_tensor_map(self):
            return fe.materials.ThermoElastic(nu=0.3, alpha=1e-3)

    thermal = Thermal(mesh, vec=1, dim=3, ele_type="HEX8")
    bc_th = fe.DirichletBCConfig([fe.DirichletBCSpec(lambda p: np.isclose(p[0], 0.0), "all", 0.0),
                                  fe.DirichletBCSpec(lambda p: np.isclose(p[0], 4.0), "all", 100.0)]).create_bc(thermal)
    opts = fe.KrylovSolverOptions(solver="cg", tol=1e-12, atol=0.0, use_jacobi_preconditioner=True)
    print("thermal solution as coefficient", flush=True)
    T = fe.create_solver(thermal, bc_th, opts, linear=True)(
        fe.TracedParams(volume_vars=(fe.TracedParams.create_cell_var(thermal, 1.0),)),
        fe.zero_like_initial_guess(thermal, bc_th))
    assert isinstance(T, fe.Solution) and abs(float(T.node_var().max()) - 100.0) < 1e-9
    mech = ThermoElastic(mesh, vec=3, dim=3, ele_type="HEX8")
    bc_me = fe.DirichletBCConfig([fe.DirichletBCSpec(lambda p: np.isclose(p[0], 0.0), "all", 0.0)]).create_bc(mech)
    u = fe.create_solver(mech, bc_me, opts, linear=True)(
        fe.TracedParams(volume_vars=(fe.TracedParams.create_cell_var(mech, 100.0), T)),
        fe.zero_like_initial_guess(mech, bc_me))
    ux = np.asarray(u).reshape(-1, 3)[:, 0]
    assert ux.max() > 1e-3
