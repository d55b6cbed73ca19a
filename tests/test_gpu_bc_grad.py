"""Gradients with respect to Dirichlet values through the linear and Newton
solvers (the cases of the reference's tests/test_bc_grad.py: a 2 x 2 QUAD4
plane-strain square, left edge clamped, right edge x fixed, one non-zero
prescribed value), checked against central differences of re-solved states;
plus the closed form of the linear solve under symmetric elimination
(d(u.u)/d bc_vals = 2 bc_vals, the free DOFs seeing the Dirichlet values
only through the initial guess) and a HEX8 Neo-Hookean Newton case."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _setup():
    import feax_amd as fe
    mesh = fe.mesh.rectangle_mesh(Nx=2, Ny=2, domain_x=1.0, domain_y=1.0, ele_type="QUAD4")

    class Elastic(fe.Problem):
        def get_tensor_map(self):
            # sigma = 2 eps + tr(eps) I: lambda = mu = 1
            return fe.materials.IsotropicLinearElastic(nu=0.25, E=2.5)

    problem = Elastic(mesh, vec=2, dim=2, ele_type="QUAD4")
    bc = fe.DirichletBCConfig([
        fe.DirichletBCSpec(location=lambda p: np.isclose(p[0], 0.0, atol=1e-6), component="all", value=0.0),
        fe.DirichletBCSpec(location=lambda p: np.isclose(p[0], 1.0, atol=1e-6), component="x", value=0.0),
    ]).create_bc(problem)
    tp = fe.TracedParams(volume_vars=())
    vals = bc.vals_numpy().copy()
    vals[-1] = 0.1
    initial = fe.zero_like_initial_guess(problem, bc)
    return problem, bc, tp, vals, initial


def _check(solver, bc, tp, vals, initial, tol):
    import torch

    def loss(v):
        with torch.no_grad():
            return float((solver(tp, initial, bc=bc.replace_vals(torch.as_tensor(v, device="cuda"))) ** 2).sum())

    bv = torch.tensor(vals, device="cuda", requires_grad=True)
    (solver(tp, initial, bc=bc.replace_vals(bv)) ** 2).sum().backward()
    g = bv.grad.detach().cpu().numpy()
    fd = np.zeros_like(vals)
    for i in range(len(vals)):
        p, m = vals.copy(), vals.copy()
        p[i] += 1e-5
        m[i] -= 1e-5
        fd[i] = (loss(p) - loss(m)) / 2e-5
    err = np.linalg.norm(g - fd) / (np.linalg.norm(fd) + 1e-30)
    print("bc grad rel err vs FD", err, flush=True)
    assert err < tol, (g, fd)
    return g, fd


def test_bc_grad_linear():
    import feax_amd as fe
    problem, bc, tp, vals, initial = _setup()
    solver = fe.create_solver(problem, bc, solver_options=fe.KrylovSolverOptions(solver="cg"), linear=True)
    g, _ = _check(solver, bc, tp, vals, initial, 1e-6)
    assert np.allclose(g, 2 * vals, rtol=0, atol=1e-14)


def test_bc_grad_linear_nonsymmetric():
    """symmetric_elimination=False keeps the Dirichlet columns: the free DOFs
    respond to bc_vals and the gradient carries -(K lambda_f)_b."""
    import feax_amd as fe
    problem, bc, tp, vals, initial = _setup()
    # feax's create_solver(linear=True) always eliminates symmetrically; the
    # row-only elimination of the linear solve is create_linear_solver's option
    solver = fe.create_linear_solver(problem, bc, fe.KrylovSolverOptions(solver="bicgstab", tol=1e-12),
                                     symmetric_elimination=False)
    g, _ = _check(solver, bc, tp, vals, initial, 1e-6)
    assert not np.allclose(g, 2 * vals, atol=1e-6)


@pytest.mark.parametrize("opts", ["cg", "direct"])
def test_bc_grad_newton(opts):
    import feax_amd as fe
    problem, bc, tp, vals, initial = _setup()
    o = fe.DirectSolverOptions(solver="cudss") if opts == "direct" else fe.KrylovSolverOptions(solver="cg")
    solver = fe.create_solver(problem, bc, solver_options=o, linear=False, newton_options=fe.NewtonOptions())
    _check(solver, bc, tp, vals, initial, 1e-6)


def test_bc_grad_newton_nonsymmetric():
    import feax_amd as fe
    problem, bc, tp, vals, initial = _setup()
    solver = fe.create_solver(problem, bc, solver_options=fe.KrylovSolverOptions(solver="gmres"), linear=False,
                              newton_options=fe.NewtonOptions(), symmetric_elimination=False)
    _check(solver, bc, tp, vals, initial, 1e-6)


def test_bc_grad_hex8_neo_hookean_newton():
    """A prescribed compression of a Neo-Hookean HEX8 block: d(u.u)/d bc_vals
    through Newton and d(u.u)/d mu in the same backward pass, against
    central differences."""
    import torch
    import feax_amd as fe
    mesh = fe.mesh.StructuredGrid((2, 2, 3)).to_mesh()

    class NH(fe.Problem):
        def get_tensor_map(self):
            return None

        def get_energy_density(self):
            return fe.materials.CompressibleNeoHookean(kappa=10.0)

    problem = NH(mesh, vec=3, dim=3, ele_type="HEX8")
    bc = fe.DirichletBCConfig([fe.DirichletBCSpec(lambda p: np.isclose(p[2], 0.0), "all", 0.0),
                               fe.DirichletBCSpec(lambda p: np.isclose(p[2], 3.0), "z", -0.15)]).create_bc(problem)
    rng = np.random.default_rng(2)
    mu0 = rng.uniform(0.5, 1.5, problem.num_cells)
    vals = bc.vals_numpy() + np.where(bc.vals_numpy() != 0, rng.uniform(-0.02, 0.02, len(bc.bc_rows)), 0.0)
    solver = fe.create_solver(problem, bc, fe.DirectSolverOptions(solver="cudss"),
                              newton_options=fe.NewtonOptions(tol=1e-12, rel_tol=1e-14), linear=False)
    ig = fe.zero_like_initial_guess(problem, bc.replace_vals(vals))

    def J(v, m):
        with torch.no_grad():
            tpv = fe.TracedParams(volume_vars=(torch.as_tensor(m, device="cuda"),))
            return float((solver(tpv, ig, bc=bc.replace_vals(torch.as_tensor(v, device="cuda"))) ** 2).sum())

    bv = torch.tensor(vals, device="cuda", requires_grad=True)
    mu = torch.tensor(mu0, device="cuda", requires_grad=True)
    (solver(fe.TracedParams(volume_vars=(mu,)), ig, bc=bc.replace_vals(bv)) ** 2).sum().backward()
    gb, gm = bv.grad.cpu().numpy(), mu.grad.cpu().numpy()
    for i in rng.choice(len(vals), 6, replace=False):
        p, m = vals.copy(), vals.copy()
        p[i] += 1e-6
        m[i] -= 1e-6
        fd = (J(p, mu0) - J(m, mu0)) / 2e-6
        assert abs(fd - gb[i]) <= 1e-6 * np.abs(gb).max(), (i, fd, gb[i])
    for e in rng.choice(len(mu0), 3, replace=False):
        p, m = mu0.copy(), mu0.copy()
        p[e] += 1e-6
        m[e] -= 1e-6
        fd = (J(vals, p) - J(vals, m)) / 2e-6
        assert abs(fd - gm[e]) <= 1e-6 * np.abs(gm).max(), (e, fd, gm[e])
