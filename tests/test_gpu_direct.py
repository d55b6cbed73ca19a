"""GPU parity of the direct solver (banded Cholesky, fx_band_*).

The factor against the matrix it factors (L L^T = P J P^T), the solution
against scipy's direct solve of the oracle system (the 50x5x5 cantilever and a box whose
bandwidth spans several factor blocks), gradients against the CG path, the
RCM ordering of a shuffled mesh, and the non-positive-definite report.
"""

import numpy as np
import pytest

from conftest import cantilever, oracle_model
from test_structure_host import shuffled_mesh

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy()


def _simp(dims=None, pts_cells=None, seed=3):
    import feax_amd as fe

    class Simp(fe.Problem):
        def get_tensor_map(self):
            return fe.materials.SIMPLinearElastic(nu=0.3, E0=1.0, Emin=1e-3, penal=3.0)

        def get_surface_maps(self):
            return [fe.materials.SurfaceTraction((0.0, 0.0, 1.0))]

    mesh = fe.StructuredGrid(dims).to_mesh() if dims is not None else fe.Mesh(*pts_cells)
    xmax = mesh.points[:, 0].max()
    prob = Simp(mesh, vec=3, dim=3, location_fns=[lambda x: np.isclose(x[0], xmax)])
    bc = fe.DirichletBCConfig([fe.DirichletBCSpec(lambda x: np.isclose(x[0], 0.0), "all", 0.0)]).create_bc(prob)
    rho = np.random.default_rng(seed).uniform(0.3, 1.0, prob.num_cells)
    return prob, bc, rho


def _tp(prob, rho, t=0.5, grad=False):
    import torch
    import feax_amd as fe
    r = torch.as_tensor(rho, device="cuda")
    if grad:
        r.requires_grad_(True)
    return fe.TracedParams(volume_vars=(r,), surface_vars=[(fe.TracedParams.create_uniform_surface_var(prob, t),)]), r


def _oracle_system(prob, bc, rho, t=0.5):
    from oracle import hex8, krylov
    m = hex8.Model(prob.points, prob.cells.astype(np.int64), prob.boundary_inds_list)
    mat = hex8.Material("simp", nu=0.3, E0=1.0, Emin=1e-3, penal=3.0)
    tr = [hex8.constant_traction(prob.boundary_inds_list[0], t)]
    _, info = krylov.linear_solve(m, mat, rho, tr, bc.bc_rows, bc.bc_vals, maxiter=1)
    return info["K"], info["b"]


@pytest.mark.parametrize("case", ["grid", "shuffled"])
def test_band_factor_reproduces_the_matrix(case):
    import torch
    import feax_amd as fe
    from feax_amd.assembler import assemble_csr
    from feax_amd.solvers.direct import BandCholesky
    if case == "grid":
        prob, bc, rho = _simp((6, 3, 4))
    else:
        prob, bc, rho = _simp(pts_cells=shuffled_mesh((6, 3, 3), 5))
    ts = fe.TracedStructure.from_problem(prob)
    tp, _ = _tp(prob, rho)
    J = assemble_csr(prob, ts, tp, None, bc, True, torch.float64)
    F = BandCholesky(prob, ts.device).factor(J)
    torch.cuda.synchronize()
    n, ldab, bw = F.n, F.ldab, F.bw
    ab = _np(F.ab[: n * ldab]).reshape(n, ldab)
    L = np.zeros((n, n))
    for j in range(n):
        m = min(bw + 1, n - j)
        L[j:j + m, j] = ab[j, :m]
    import scipy.sparse as sp
    K = sp.csr_matrix((_np(J.data), _np(J.indices), _np(J.indptr)), shape=(n, n)).toarray()
    perm = np.arange(n) if F.perm is None else _np(F.perm)
    PK = K[np.ix_(perm, perm)]
    err = np.abs(L @ L.T - PK).max() / np.abs(K).max()
    print(case, "n", n, "bw", bw, "ordering", "natural" if F.perm is None else "rcm", "|LL^T - PKP^T|", err)
    assert err < 1e-13
    assert (case == "shuffled") == (F.perm is not None)


@pytest.mark.parametrize("case", ["cantilever50x5x5", "grid12x8x8"])
def test_direct_solve_matches_scipy_direct_solve(case):
    import scipy.sparse.linalg as spla
    import feax_amd as fe
    if case == "cantilever50x5x5":
        prob, bc, tp = cantilever((50, 5, 5), (2.0, 2.0, 2.0), 100.0)
        from oracle import hex8, krylov
        m = oracle_model(prob)
        mat = hex8.Material("linear_elastic", nu=0.3)
        _, info = krylov.linear_solve(m, mat, np.full(m.N, 70e3),
                                      [hex8.constant_traction(prob.boundary_inds_list[0], 1e-3)],
                                      bc.bc_rows, bc.bc_vals, maxiter=1)
        K, b = info["K"], info["b"]
    else:
        prob, bc, rho = _simp((12, 8, 8))
        tp, _ = _tp(prob, rho)
        K, b = _oracle_system(prob, bc, rho)
    solver = fe.create_solver(prob, bc, fe.DirectSolverOptions(), linear=True, traced_params=tp)
    u = np.asarray(solver(tp))
    ud = spla.spsolve(K.tocsc(), b)
    err = np.linalg.norm(u - ud) / np.linalg.norm(ud)
    res = np.linalg.norm(K @ u - b) / np.linalg.norm(b)
    res_d = np.linalg.norm(K @ ud - b) / np.linalg.norm(b)
    print(case, "rel err vs scipy", err, "relres", res, "scipy's relres", res_d)
    # both are backward-stable direct solves: residuals at the rounding
    # level of the (ill-conditioned, slender 50x5x5) system, errors ~ cond * eps
    assert err < 1e-9 and res < 100 * res_d + 1e-14
    # a second call refactors cleanly; create_solver's default options are
    # this direct solver, as feax's are (feax/solver.py:293-301)
    u2 = np.asarray(solver(tp))
    assert np.array_equal(u, u2)
    u3 = np.asarray(fe.create_solver(prob, bc, linear=True, traced_params=tp)(tp))
    assert np.array_equal(u, u3)


def test_direct_gradient_matches_cg_gradient():
    import feax_amd as fe
    prob, bc, rho = _simp((8, 4, 4), seed=6)
    grads = []
    for opts in (fe.DirectSolverOptions(solver="cudss"), fe.KrylovSolverOptions(solver="cg", tol=1e-13, atol=0.0,
                                                                  use_jacobi_preconditioner=True)):
        solver = fe.create_solver(prob, bc, opts, linear=True, return_solution=False)
        tp, r = _tp(prob, rho, grad=True)
        u = solver(tp)
        (u * u).sum().backward()
        grads.append(_np(r.grad))
    err = np.abs(grads[0] - grads[1]).max() / np.abs(grads[1]).max()
    print("gradient rel diff direct vs cg", err)
    assert err < 1e-9
    # mixed: direct forward, CG adjoint; and CG forward, direct adjoint
    for fwd, adj in ((fe.DirectSolverOptions(solver="cudss"), fe.KrylovSolverOptions(solver="cg", tol=1e-13, atol=0.0,
                                                                        use_jacobi_preconditioner=True)),
                     (fe.KrylovSolverOptions(solver="cg", tol=1e-13, atol=0.0, use_jacobi_preconditioner=True),
                      fe.DirectSolverOptions(solver="cudss"))):
        solver = fe.create_solver(prob, bc, fwd, adjoint_solver_options=adj, linear=True, return_solution=False)
        tp, r = _tp(prob, rho, grad=True)
        (solver(tp) ** 2).sum().backward()
        assert np.abs(_np(r.grad) - grads[1]).max() <= 1e-9 * np.abs(grads[1]).max()


def test_indefinite_matrices_take_ldlt_and_singular_ones_are_reported():
    """A symmetric indefinite matrix (J - sigma I, and -J) fails Cholesky and
    is factored as L D L^T without pivoting (solution vs scipy); a matrix
    with an empty free row and column has an exactly zero pivot."""
    import torch
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    import feax_amd as fe
    from feax_amd.assembler import assemble_csr
    from feax_amd.solvers.direct import BandCholesky
    prob, bc, rho = _simp((6, 3, 3))
    ts = fe.TracedStructure.from_problem(prob)
    tp, _ = _tp(prob, rho)
    J = assemble_csr(prob, ts, tp, None, bc, True, torch.float64)
    n = ts.n
    K = sp.csr_matrix((_np(J.data), _np(J.indices), _np(J.indptr)), shape=(n, n))
    rows = np.repeat(np.arange(n), np.diff(K.indptr))
    sigma = 0.37 * np.median(K.diagonal())
    data_shift = np.where(rows == K.indices, K.data - sigma, K.data)
    b = torch.as_tensor(np.cos(np.arange(n)), device="cuda")
    for data in (data_shift, -K.data):
        A = J.with_data(torch.as_tensor(data, device="cuda"))
        F = BandCholesky(prob, ts.device).factor(A)
        assert F.ldlt
        x = _np(F.solve(b))
        Ks = sp.csr_matrix((data, K.indices, K.indptr), shape=(n, n))
        xd = spla.spsolve(Ks.tocsc(), _np(b))
        err = np.linalg.norm(x - xd) / np.linalg.norm(xd)
        print("ldlt", "rel err vs scipy", err, "negative eigenvalues",
              int((np.linalg.eigvalsh(Ks.toarray()) < 0).sum()))
        assert err < 1e-9
    free = int(np.nonzero(~bc.bc_mask)[0][5])
    empty = np.where((rows == free) | (K.indices == free), 0.0, K.data)
    with pytest.raises(RuntimeError, match="zero pivot"):
        BandCholesky(prob, ts.device).factor(J.with_data(torch.as_tensor(empty, device="cuda")))
    # the SPD matrix still takes Cholesky, and solves after the failures
    F = BandCholesky(prob, ts.device).factor(J)
    assert not F.ldlt
    x = F.solve(b)
    assert np.linalg.norm(K @ _np(x) - _np(b)) <= 1e-10 * np.linalg.norm(_np(b))


@pytest.mark.parametrize("dims", [(30, 10, 10), (9, 4, 3)])
def test_blocked_inverse_solve_matches_the_sub_block_solve_and_scipy(dims):
    """fx_band_solve_inv_f64 (inverse diagonal blocks, two launches per block)
    against fx_band_solve_f64 and scipy, on a box whose band spans many
    512-row solve blocks and on one small enough for a single block; the
    L D L^T form (unit L) on an indefinite shift of the same matrix."""
    import torch
    import scipy.sparse.linalg as spl
    import feax_amd as fe
    from feax_amd import _native as nat
    from feax_amd.assembler import assemble_csr, residual
    from feax_amd.solvers.direct import BandCholesky
    prob, bc, rho = _simp(dims)
    ts = fe.TracedStructure.from_problem(prob)
    tp, _ = _tp(prob, rho)
    print("case: blocked-inverse solve", dims, flush=True)
    J = assemble_csr(prob, ts, tp, None, bc, True, torch.float64)
    F = BandCholesky(prob, ts.device).factor(J)
    nbs = int(nat.lib().fx_band_inv_block(F.bw))
    print("n", F.n, "bw", F.bw, "nbs", nbs, "blocks", (F.n + nbs - 1) // nbs)
    b = residual(prob, ts, tp, torch.zeros(ts.n, dtype=torch.float64, device="cuda"), bc, negate=True)
    x1 = F.solve(b)
    x2 = F.solve(b, blocked_inverse=False)
    import scipy.sparse as sp
    K = sp.csr_matrix((_np(J.data), _np(J.indices), _np(J.indptr)), shape=(ts.n, ts.n))
    xs = spl.spsolve(K.tocsc(), _np(b))
    scale = np.abs(xs).max()
    print("inv vs sub-block", np.abs(_np(x1) - _np(x2)).max() / scale, "inv vs scipy",
          np.abs(_np(x1) - xs).max() / scale)
    assert np.abs(_np(x1) - _np(x2)).max() <= 1e-10 * scale
    assert np.abs(_np(x1) - xs).max() <= 1e-9 * scale
    # an indefinite matrix (J - sigma I on the free rows) takes L D L^T
    sigma = 0.5 * float(spl.eigsh(K, k=1, which="LA", return_eigenvectors=False)[0])
    d = _np(J.data).copy()
    rows = np.repeat(np.arange(ts.n), np.diff(_np(J.indptr)))
    diag = (_np(J.indices) == rows) & ~np.isin(rows, bc.bc_rows)
    d[diag] -= sigma
    A = J.with_data(torch.as_tensor(d, device="cuda"))
    G = BandCholesky(prob, ts.device).factor(A)
    assert G.ldlt
    y1 = G.solve(b)
    y2 = G.solve(b, blocked_inverse=False)
    Ks = sp.csr_matrix((d, _np(J.indices), _np(J.indptr)), shape=(ts.n, ts.n))
    ys = spl.spsolve(Ks.tocsc(), _np(b))
    sc = np.abs(ys).max()
    print("ldlt inv vs sub-block", np.abs(_np(y1) - _np(y2)).max() / sc, "vs scipy", np.abs(_np(y1) - ys).max() / sc)
    assert np.abs(_np(y1) - _np(y2)).max() <= 1e-8 * sc
    assert np.abs(_np(y1) - ys).max() <= 1e-7 * sc
