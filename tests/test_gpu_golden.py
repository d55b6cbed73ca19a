"""GPU results against the committed golden fixtures (tests/golden/)."""

import os

import numpy as np
import pytest

from conftest import cantilever

pytestmark = pytest.mark.gpu
G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
EPS64 = np.finfo(np.float64).eps


def _load(name):
    return np.load(os.path.join(G, name), allow_pickle=False)


def test_fixture_2x2x2_assembly_residual_solve():
    import torch
    import feax_amd as fe
    from feax_amd.assembler import assemble_csr, residual
    f = _load("fixture_2x2x2.npz")
    p, bc, tp = cantilever((2, 2, 2), (1.0, 1.0, 1.0), 2.0)
    ts = fe.TracedStructure.from_problem(p)
    A = assemble_csr(p, ts, tp, None, bc, True, torch.float64)
    data = A.data.cpu().numpy()
    rowmax = np.repeat([np.abs(f["data_bc"][a:b]).max() for a, b in zip(f["indptr"][:-1], f["indptr"][1:])],
                       np.diff(f["indptr"]))
    assert np.all(np.abs(data - f["data_bc"]) <= 64 * EPS64 * rowmax)
    r0 = residual(p, ts, tp, torch.zeros(p.num_total_dofs_all_vars, dtype=torch.float64, device="cuda"))
    np.testing.assert_allclose(r0.cpu().numpy(), f["res0"], rtol=0, atol=1e-14 * np.abs(f["res0"]).max())
    opts = fe.KrylovSolverOptions(solver="cg", tol=1e-12, atol=0.0, maxiter=1000, use_jacobi_preconditioner=True)
    u = np.asarray(fe.create_solver(p, bc, opts, linear=True)(tp))
    assert np.linalg.norm(u - f["u"]) / np.linalg.norm(f["u"]) < 1e-10


def test_c1_sampled_values_and_solution():
    import torch
    import feax_amd as fe
    from feax_amd.assembler import assemble_csr
    f = _load("cantilever_50x5x5_samples.npz")
    p, bc, tp = cantilever((50, 5, 5), (2.0, 2.0, 2.0), 100.0)
    ts = fe.TracedStructure.from_problem(p)
    data = assemble_csr(p, ts, tp, None, bc, True, torch.float64).data.cpu().numpy()
    vals = f["values"]
    assert np.abs(data[f["slots"]] - vals).max() <= 1e-13 * np.abs(vals).max()
    for precision in ("fp64", "mixed"):
        opts = fe.KrylovSolverOptions(solver="cg", tol=1e-10, atol=0.0, maxiter=10000,
                                      use_jacobi_preconditioner=True, precision=precision)
        u = np.asarray(fe.create_solver(p, bc, opts, linear=True)(tp))[f["dofs"]]
        assert np.abs(u - f["u"]).max() <= 1e-8 * np.abs(f["u"]).max(), precision


def test_simp_fixture_gradient():
    """dC/drho of the 4x2x2 SIMP fixture via torch.autograd through the HIP
    solve, and via the fused self-adjoint kernel."""
    import torch
    import feax_amd as fe
    from feax_amd.gene import compliance_and_sensitivity, create_compliance_fn
    f = _load("simp_4x2x2.npz")

    class Simp(fe.Problem):
        def get_tensor_map(self):
            return fe.materials.SIMPLinearElastic(nu=0.3, E0=1.0, Emin=1e-9, penal=3.0)

        def get_surface_maps(self):
            return [fe.materials.SurfaceTraction((0.0, 0.0, 1.0))]

    p = Simp(fe.StructuredGrid((4, 2, 2)).to_mesh(), vec=3, dim=3, location_fns=[lambda x: np.isclose(x[0], 4.0)])
    bc = fe.DirichletBCConfig([fe.DirichletBCSpec(lambda x: np.isclose(x[0], 0.0), "all", 0.0)]).create_bc(p)
    t = fe.TracedParams.create_uniform_surface_var(p, 0.25)
    rho = torch.tensor(f["rho"], device="cuda", requires_grad=True)
    opts = fe.KrylovSolverOptions(solver="cg", tol=1e-12, atol=0.0, maxiter=1000, use_jacobi_preconditioner=True)
    solver = fe.create_solver(p, bc, opts, linear=True, return_solution=False)
    comp = create_compliance_fn(p, 0.25)
    u = solver(fe.TracedParams(volume_vars=(rho,), surface_vars=[(t,)]))
    C = comp(u)
    C.backward()
    assert abs(float(C.detach()) - float(f["compliance"])) <= 1e-10 * float(f["compliance"])
    np.testing.assert_allclose(rho.grad.cpu().numpy(), f["dc"], rtol=1e-9, atol=1e-12 * np.abs(f["dc"]).max())
    _, dc = compliance_and_sensitivity(p, fe.TracedParams(volume_vars=(f["rho"],)), u.detach(), 0.25)
    np.testing.assert_allclose(dc.cpu().numpy(), f["dc"], rtol=1e-9, atol=1e-12 * np.abs(f["dc"]).max())


@pytest.mark.parametrize("pre", ["jacobi", "gmg-fp64", "gmg-mixed"])
def test_c2_full_size_solution_matches_oracle_samples(pre):
    """200x40x40 (1,013,643 DOFs): sampled DOFs of the GPU solve vs the
    oracle's Jacobi-PCG to rel 1e-12 (cantilever_200x40x40_oracle.npz)."""
    import feax_amd as fe
    f = _load("cantilever_200x40x40_oracle.npz")
    p, bc, tp = cantilever((200, 40, 40), (1.0, 1.0, 1.0), 200.0)
    mg = fe.GeometricMultigrid() if pre.startswith("gmg") else None
    opts = fe.KrylovSolverOptions(solver="cg", tol=1e-11, atol=0.0, maxiter=20000, preconditioner=mg,
                                  use_jacobi_preconditioner=mg is None,
                                  precision="mixed" if pre == "gmg-mixed" else "fp64")
    solver = fe.create_solver(p, bc, opts, linear=True, return_solution=False)
    u = solver(tp).cpu().numpy()
    err = np.abs(u[f["dofs"]] - f["u"]).max() / float(f["u_max"])
    print(pre, "iters", solver.stats.iters, "max sampled err / max|u|", err)
    assert err < 1e-8
    assert abs(np.abs(u).max() - float(f["u_max"])) <= 1e-8 * float(f["u_max"])


def test_c3_full_size_compliance_sensitivity_matches_oracle_samples():
    """120x40x20 SIMP cantilever: compliance and the fused dC/drho kernel
    after a GMG solve vs the oracle's closed form (simp_120x40x20_oracle.npz)."""
    import torch
    import feax_amd as fe
    from feax_amd import gene
    f = _load("simp_120x40x20_oracle.npz")
    dims = (120, 40, 20)

    class Simp(fe.Problem):
        def get_tensor_map(self):
            return fe.materials.SIMPLinearElastic(nu=0.3, E0=1.0, Emin=1e-9, penal=3.0)

        def get_surface_maps(self):
            return [fe.materials.SurfaceTraction((0.0, 0.0, 1.0))]

    patch = [lambda x: np.isclose(x[0], 120) & (np.abs(x[1] - 20) <= 1 + 1e-9) & (np.abs(x[2] - 10) <= 1 + 1e-9)]
    prob = Simp(fe.StructuredGrid(dims).to_mesh(), vec=3, dim=3, location_fns=patch)
    bc = fe.DirichletBCConfig([fe.DirichletBCSpec(lambda x: np.isclose(x[0], 0.0), "all", 0.0)]).create_bc(prob)
    rho = np.random.default_rng(20261203).uniform(0.3, 1.0, prob.num_cells)
    t = fe.TracedParams.create_uniform_surface_var(prob, 0.25)
    tp = fe.TracedParams(volume_vars=(torch.as_tensor(rho, device="cuda"),), surface_vars=[(t,)])
    opts = fe.KrylovSolverOptions(solver="cg", tol=1e-11, atol=0.0, preconditioner=fe.GeometricMultigrid())
    solver = fe.create_solver(prob, bc, opts, linear=True, return_solution=False)
    u = solver(tp)
    C, dc = gene.compliance_and_sensitivity(prob, tp, u, surface_load_params=0.25)
    dc = dc.cpu().numpy()
    err = np.abs(dc[f["cells"]] - f["dc"]).max() / float(f["dc_max"])
    print("simp 120x40x20 iters", solver.stats.iters, "dc err", err, "C", float(C), float(f["compliance"]))
    assert err < 1e-7
    assert abs(float(C) - float(f["compliance"])) <= 1e-9 * abs(float(f["compliance"]))
