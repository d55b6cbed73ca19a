"""GPU parity of the geometric-multigrid preconditioner.

The V-cycle (fx_gmg_vcycle) and the Lanczos spectrum bounds against the
oracle's V-cycle on the oracle's own hierarchy (oracle/gmg.py, pinned by
the Galerkin identity, the Chebyshev closed form and SPD-ness); the
GMG-PCG solve against scipy's direct solve of the oracle system; gradients
through the GMG solver against the Jacobi path.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy()


def _simp_problem(dims, seed=5, load=True):
    import feax_amd as fe
    nx, ny, nz = dims

    class Simp(fe.Problem):
        def get_tensor_map(self):
            return fe.materials.SIMPLinearElastic(nu=0.3, E0=1.0, Emin=1e-3, penal=3.0)

        def get_surface_maps(self):
            return [fe.materials.SurfaceTraction((0.0, 0.0, 1.0))]

    mesh = fe.StructuredGrid(dims).to_mesh()
    locs = [lambda p: np.isclose(p[0], nx)] if load else []
    prob = Simp(mesh, vec=3, dim=3, location_fns=locs)
    bc = fe.DirichletBCConfig([fe.DirichletBCSpec(lambda p: np.isclose(p[0], 0.0), "all", 0.0)]).create_bc(prob)
    rho = np.random.default_rng(seed).uniform(0.2, 1.0, prob.num_cells)
    return prob, bc, rho


def _hierarchy(prob, bc, rho, mg, fp32=False):
    import torch
    import feax_amd as fe
    from feax_amd.assembler import assemble_csr
    from feax_amd.solvers.gmg import GMGHierarchy
    ts = fe.TracedStructure.from_problem(prob)
    tp = fe.TracedParams(volume_vars=(torch.as_tensor(rho, device="cuda"),))
    u0 = torch.zeros(ts.n, dtype=torch.float64, device="cuda")
    J, dinv = assemble_csr(prob, ts, tp, u0, bc, True, torch.float64, jacobi_shift=1e-12)
    return GMGHierarchy(prob, ts, tp, u0, bc, J, dinv, mg, fp32=fp32), ts


def _oracle_levels(prob, bc, rho, mg, dense=True):
    from oracle import gmg, hex8
    from feax_amd.solvers.gmg import structured_dims
    dims, _ = structured_dims(prob)
    mat = hex8.Material("simp", nu=0.3, E0=1.0, Emin=1e-3, penal=3.0)
    return gmg.hierarchy(dims, rho, mat, bc.bc_mask, min_cells=mg.min_cells, lmax_iters=mg.lmax_iters,
                         nlev=mg.max_levels, dense=dense, coarse_dofs=mg.coarse_dofs)


@pytest.mark.parametrize("dims,coarse", [((8, 4, 4), "dense"), ((8, 4, 6), "dense"), ((8, 4, 4), "chebyshev"),
                                         ((6, 2, 2), "chebyshev"), ((10, 6, 6), "dense"), ((5, 3, 7), "chebyshev")])
def test_vcycle_matches_oracle(dims, coarse):
    import torch
    from oracle import gmg
    from feax_amd.solvers.gmg import GeometricMultigrid
    prob, bc, rho = _simp_problem(dims)
    mg = GeometricMultigrid(coarse=coarse, min_cells=1, coarse_degree=20, coarse_dofs=1)
    H, ts = _hierarchy(prob, bc, rho, mg)
    levels = _oracle_levels(prob, bc, rho, mg, dense=coarse == "dense")
    assert H.nlev == len(levels) and [tuple(d) for d in H.S.dims] == [l["dims"] for l in levels]
    for l, lev in enumerate(levels):
        if lev["lmax"] > 0:
            assert abs(H.lmax[l] - lev["lmax"]) <= 1e-9 * lev["lmax"], (l, H.lmax[l], lev["lmax"])
        if lev["lmin"] > 0:
            assert abs(H.lmin[l] - lev["lmin"]) <= 1e-8 * lev["lmax"], (l, H.lmin[l], lev["lmin"])
    rng = np.random.default_rng(7)
    r = np.where(bc.bc_mask, 0.0, rng.standard_normal(ts.n))
    z = _np(H.vcycle(torch.as_tensor(r, device="cuda")))
    ref = gmg.vcycle(levels, r, mg.smooth_degree, mg.coarse_degree, mg.cheb_ratio)
    err = np.abs(z - ref).max() / np.abs(ref).max()
    print(dims, coarse, "levels", H.S.dims, "vcycle rel err", err)
    assert err < 1e-9
    assert np.all(z[bc.bc_mask] == 0.0)


def test_fp32_vcycle_close_to_oracle():
    import torch
    from oracle import gmg
    from feax_amd.solvers.gmg import GeometricMultigrid
    prob, bc, rho = _simp_problem((8, 4, 4))
    mg = GeometricMultigrid(min_cells=1, coarse_dofs=1)
    H, ts = _hierarchy(prob, bc, rho, mg, fp32=True)
    levels = _oracle_levels(prob, bc, rho, mg)
    r = np.where(bc.bc_mask, 0.0, np.random.default_rng(8).standard_normal(ts.n))
    z = _np(H.vcycle(torch.as_tensor(r, device="cuda")))
    ref = gmg.vcycle(levels, r, mg.smooth_degree, mg.coarse_degree, mg.cheb_ratio)
    err = np.abs(z - ref).max() / np.abs(ref).max()
    print("fp32 vcycle rel err", err)
    assert err < 1e-4


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_gmg_solve_matches_direct_solve(precision):
    import torch
    import feax_amd as fe
    from oracle import hex8, krylov
    import scipy.sparse.linalg as spla
    from feax_amd.solvers.gmg import GeometricMultigrid
    prob, bc, rho = _simp_problem((16, 8, 8), seed=9)
    opts = fe.KrylovSolverOptions(solver="cg", tol=1e-10, atol=0.0, precision=precision,
                                  preconditioner=GeometricMultigrid(coarse_dofs=200))
    solver = fe.create_solver(prob, bc, opts, linear=True)
    t = fe.TracedParams.create_uniform_surface_var(prob, 0.5)
    tp = fe.TracedParams(volume_vars=(torch.as_tensor(rho, device="cuda"),), surface_vars=[(t,)])
    u = _np(solver(tp).dofs)
    m = hex8.Model(prob.points, prob.cells.astype(np.int64), prob.boundary_inds_list)
    mat = hex8.Material("simp", nu=0.3, E0=1.0, Emin=1e-3, penal=3.0)
    tr = [hex8.constant_traction(prob.boundary_inds_list[0], 0.5)]
    _, info = krylov.linear_solve(m, mat, rho, tr, bc.bc_rows, bc.bc_vals, maxiter=1)
    ud = spla.spsolve(info["K"].tocsc(), info["b"])
    err = np.linalg.norm(u - ud) / np.linalg.norm(ud)
    st = solver.stats
    print(precision, "iters", st.iters, "rel err vs direct", err)
    assert err < 1e-8
    assert st.iters < 40


def test_graph_and_direct_launches_agree_bitwise():
    import torch
    from feax_amd.solvers.gmg import GeometricMultigrid
    prob, bc, rho = _simp_problem((8, 4, 4))
    H, ts = _hierarchy(prob, bc, rho, GeometricMultigrid(min_cells=1, check_every=3, coarse_dofs=1))
    b = torch.as_tensor(np.where(bc.bc_mask, 0.0, np.random.default_rng(2).standard_normal(ts.n)), device="cuda")
    x1, i1 = H.solve(b, None, 1e-10, 0.0, 200, graph=True)
    x2, i2 = H.solve(b, None, 1e-10, 0.0, 200, graph=False)
    assert i1.iters == i2.iters and i1.iters > 0
    assert torch.equal(x1, x2)
    # maxiter stops the chunked loop exactly
    x3, i3 = H.solve(b, None, 1e-14, 0.0, 5)
    assert i3.iters == 5


def test_gradient_through_gmg_solver_matches_jacobi_path():
    import torch
    import feax_amd as fe
    from feax_amd.solvers.gmg import GeometricMultigrid
    prob, bc, rho = _simp_problem((8, 4, 4), seed=4)
    t = fe.TracedParams.create_uniform_surface_var(prob, 0.5)
    grads = []
    for pre in (GeometricMultigrid(min_cells=1, coarse_dofs=1), None):
        opts = fe.KrylovSolverOptions(solver="cg", tol=1e-12, atol=0.0, preconditioner=pre,
                                      use_jacobi_preconditioner=pre is None)
        solver = fe.create_solver(prob, bc, opts, linear=True, return_solution=False)
        r = torch.as_tensor(rho, device="cuda").requires_grad_(True)
        u = solver(fe.TracedParams(volume_vars=(r,), surface_vars=[(t,)]))
        (u * u).sum().backward()
        grads.append(_np(r.grad))
    err = np.abs(grads[0] - grads[1]).max() / np.abs(grads[1]).max()
    print("gradient rel diff gmg vs jacobi", err)
    assert err < 1e-7


def test_gmg_rejects_unstructured_mesh():
    import torch
    import feax_amd as fe
    from feax_amd.solvers.gmg import GeometricMultigrid
    prob, bc, rho = _simp_problem((4, 2, 2))
    mesh = fe.Mesh(prob.points[::-1].copy(), (len(prob.points) - 1 - prob.cells))
    prob2 = type(prob)(mesh, vec=3, dim=3)
    bc2 = fe.DirichletBCConfig([fe.DirichletBCSpec(lambda p: np.isclose(p[0], 0.0), "all", 0.0)]).create_bc(prob2)
    solver = fe.create_solver(prob2, bc2, fe.KrylovSolverOptions(solver="cg", preconditioner=GeometricMultigrid()),
                              linear=True)
    with pytest.raises(NotImplementedError, match="structured box"):
        solver(fe.TracedParams(volume_vars=(torch.as_tensor(rho, device="cuda"),)))
