"""A solve whose boundary sets and parameter field come from user functions
that reduce over their point argument (np.linalg.norm, np.all), as feax
evaluates them with jax.vmap (feax/DCboundary.py:457-470,
feax/fe.py:288-301, feax/traced_params.py:234-266).

The HIP solve is compared with a scipy direct solve of the oracle's system,
whose boundary sets and fields are built by explicit per-point loops.
"""

import numpy as np
import pytest
import scipy.sparse.linalg as spl

pytestmark = pytest.mark.gpu

L, W = 10.0, 6.0
CENTER = np.array([L, W / 2, W / 2])
E0 = 70e3


def _patch(p):
    """Spherical Dirichlet patch on the right face: |p - c| < 2.1."""
    return np.isclose(p[0], L) & (np.linalg.norm(p - CENTER) < 2.1)


def _patch_value(p):
    """Prescribed u_x on the patch, a reduction over the point."""
    return 1e-3 * (2.1 - np.linalg.norm(p - CENTER))


def _loaded(p):
    """Top face z = W restricted to x <= 4 by np.all over the point."""
    return np.isclose(p[2], W) & np.all(p[:1] <= 4.0)


def _modulus(p):
    return E0 * (1.0 + 0.05 * np.linalg.norm(p) / L)


def test_spherical_dirichlet_patch_solve():
    import feax_amd as fe
    from oracle import hex8, krylov, structure

    class LinearElasticity(fe.Problem):
        def get_tensor_map(self):
            return fe.materials.IsotropicLinearElastic(nu=0.3)

        def get_surface_maps(self):
            return [fe.materials.SurfaceTraction((0.0, 0.0, 1.0))]

    dims = (10, 6, 6)
    mesh = fe.StructuredGrid(dims).to_mesh()
    problem = LinearElasticity(mesh, vec=3, dim=3, location_fns=[_loaded])
    bc = fe.DirichletBCConfig([fe.DirichletBCSpec(lambda p: np.isclose(p[0], 0.0), "all", 0.0),
                               fe.DirichletBCSpec(_patch, "x", _patch_value)]).create_bc(problem)
    E = fe.TracedParams.create_node_var_from_fn(problem, _modulus)
    t = fe.TracedParams.create_uniform_surface_var(problem, -2.0)
    tp = fe.TracedParams(volume_vars=(E,), surface_vars=[(t,)])

    # oracle: the same sets by explicit per-point loops
    pts, cells = problem.points, problem.cells.astype(np.int64)
    patch_nodes = [k for k, q in enumerate(pts) if _patch(q)]
    assert len(patch_nodes) == 13                      # the disc of radius 2.1 on a unit grid
    faces = structure.boundary_faces(pts, cells, _loaded)
    assert len(faces) == 4 * 6 and np.array_equal(problem.boundary_inds_list[0], faces)
    rows, vals = structure.dirichlet_rows(pts, [(lambda q: np.isclose(q[0], 0.0), "all", 0.0),
                                                (_patch, 0, _patch_value)])
    assert np.array_equal(bc.bc_rows, rows) and np.allclose(bc.bc_vals, vals, rtol=1e-15, atol=0)
    Eref = np.array([_modulus(q) for q in pts])
    assert np.allclose(E, Eref, rtol=1e-15, atol=0)

    m = hex8.Model(pts, cells, [faces])
    mat = hex8.Material("linear_elastic", nu=0.3)
    tr = [hex8.constant_traction(faces, -2.0)]
    _, info = krylov.linear_solve(m, mat, Eref, tr, rows, vals, tol=1e-12, atol=0.0)
    u0 = np.zeros(m.n)
    u0[rows] = vals
    u_dir = u0 + spl.spsolve(info["K"].tocsc(), info["b"])

    opts = fe.KrylovSolverOptions(solver="cg", tol=1e-11, atol=0.0, maxiter=20000, use_jacobi_preconditioner=True)
    solver = fe.create_solver(problem, bc, opts, linear=True, traced_params=tp)
    print("case: spherical Dirichlet patch, 10x6x6, fp64 Jacobi-PCG", flush=True)
    u = np.asarray(solver(tp, fe.zero_like_initial_guess(problem, bc)))
    err = np.linalg.norm(u - u_dir) / np.linalg.norm(u_dir)
    print(f"patch solve: iters={solver.stats.iters} rel err vs direct {err:.2e}")
    assert np.allclose(u[rows], vals, rtol=0, atol=1e-14)
    assert err < 1e-9
