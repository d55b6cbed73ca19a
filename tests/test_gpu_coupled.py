"""GPU parity of the monolithic (u, T) thermo-elastic problem (fx_hex8ut_*)
against the oracle (oracle/coupled.py): CSR assembly with the Dirichlet
elimination, the residual with traction on u and flux on T, the linear solve
through create_solver ("auto" resolves GMRES for the non-symmetric
Jacobian) against scipy's direct solve of the oracle system, and gradients
with respect to E and the surface loads against central differences."""

import numpy as np
import pytest
import scipy.sparse.linalg as spla

pytestmark = pytest.mark.gpu

LAW = (0.3, 1.2e-2, 0.5, 2.5)       # nu, alpha, t_ref, k
EPS64 = np.finfo(np.float64).eps


def _np(t):
    return t.detach().cpu().numpy()


def _setup(dims=(5, 2, 3), distort=0.08, seed=2, E_cell=True):
    import feax_amd as fe
    from oracle import structure as st
    pts, cells = st.structured_grid(dims)
    rng = np.random.default_rng(seed)
    inner = np.all((pts > 1e-9) & (pts < np.array(dims) - 1e-9), axis=1)
    pts = pts + rng.uniform(-distort, distort, pts.shape) * inner[:, None]
    nx, nz = dims[0], dims[2]

    class UT(fe.Problem):
        def get_weak_form(self):
            nu, alpha, t_ref, k = LAW
            return fe.materials.ThermoElasticCoupled(nu=nu, alpha=alpha, t_ref=t_ref, k=k)

        def get_surface_maps(self):
            return [fe.materials.SurfaceTraction((0.0, 0.0, 1.0)), fe.materials.SurfaceFlux(1.0)]

    mesh = fe.Mesh(pts, cells)
    prob = UT(mesh=[mesh, mesh], vec=[3, 1], dim=3, ele_type=["HEX8", "HEX8"],
              location_fns=[lambda p: np.isclose(p[0], nx), lambda p: np.isclose(p[2], nz)])
    left = lambda p: np.isclose(p[0], 0.0)
    bc = fe.DirichletBCConfig([fe.DirichletBCSpec(left, "all", 0.0, variable_index=0),
                               fe.DirichletBCSpec(left, 0, 1.5, variable_index=1),
                               fe.DirichletBCSpec(lambda p: np.isclose(p[2], 0.0), 0, lambda p: 0.2 * p[0],
                                                  variable_index=1)]).create_bc(prob)
    E = rng.uniform(1.0, 2.0, prob.num_cells if E_cell else prob.num_total_nodes)
    tz = rng.uniform(0.1, 0.3, (len(prob.boundary_inds_list[0]), 4))
    qz = rng.uniform(0.5, 1.0, (len(prob.boundary_inds_list[1]), 4))
    return prob, bc, E, tz, qz


def _oracle(prob, E, tz, qz):
    from oracle import coupled
    m = coupled.Model(prob.points, prob.cells, prob.boundary_inds_list)
    loads = [(0, (0.0, 0.0, 1.0), tz), (1, None, qz)]
    return m, loads


@pytest.mark.parametrize("E_cell", [True, False])
def test_coupled_assembly_and_residual_match_oracle(E_cell):
    import torch
    import feax_amd as fe
    from feax_amd.assembler import assemble_csr, residual
    from oracle import hex8
    prob, bc, E, tz, qz = _setup(E_cell=E_cell)
    m, loads = _oracle(prob, E, tz, qz)
    ts = fe.TracedStructure.from_problem(prob)
    tp = fe.TracedParams(volume_vars=(E,), surface_vars=[(tz,), (qz,)])
    print("case: coupled (u, T) assembly, E", "cell" if E_cell else "node", flush=True)
    A, dinv = assemble_csr(prob, ts, tp, None, bc, True, torch.float64, jacobi_shift=1e-12)
    assert np.array_equal(_np(A.indptr), m.csr["indptr"]) and np.array_equal(_np(A.indices), m.csr["indices"])
    ref = hex8.apply_bc_csr(m.csr_values(E, LAW), m.csr, bc.bc_mask, bc.bc_rows, True)
    K = m.scipy_csr(ref)
    scale = np.repeat(np.abs(K).max(axis=1).toarray().ravel(), np.diff(K.indptr))
    err = np.abs(_np(A.data) - ref) / scale
    print("coupled assembly max err / row max", err.max())
    assert err.max() <= 64 * EPS64
    assert np.allclose(_np(dinv), 1.0 / (K.diagonal() + 1e-12), rtol=1e-14)
    w = 0.05 * np.random.default_rng(5).standard_normal(m.n)
    r = _np(residual(prob, ts, tp, torch.as_tensor(w, device="cuda"), bc))
    rref = hex8.apply_bc_res(m.residual(w, E, LAW, loads), w, bc.bc_rows, bc.bc_vals)
    rerr = np.abs(r - rref).max() / np.abs(rref).max()
    print("coupled residual rel err", rerr)
    assert rerr <= 1e-13


def test_coupled_solve_auto_gmres_matches_direct(capsys):
    import feax_amd as fe
    from oracle import hex8
    prob, bc, E, tz, qz = _setup()
    m, loads = _oracle(prob, E, tz, qz)
    tp = fe.TracedParams(volume_vars=(E,), surface_vars=[(tz,), (qz,)])
    print("case: coupled (u, T) solve, auto", flush=True)
    opts = fe.KrylovSolverOptions(tol=1e-12, atol=0.0, restart=60, use_jacobi_preconditioner=True)
    solver = fe.create_solver(prob, bc, opts, linear=True, traced_params=tp)
    assert "matrix_property=GENERAL -> gmres" in capsys.readouterr().out
    sol = solver(tp)
    u, T = _np(sol.field(0)), _np(sol.field(1))
    w0 = np.zeros(m.n)
    w0[bc.bc_rows] = bc.bc_vals
    K = m.scipy_csr(hex8.apply_bc_csr(m.csr_values(E, LAW), m.csr, bc.bc_mask, bc.bc_rows, True))
    b = -hex8.apply_bc_res(m.residual(w0, E, LAW, loads), w0, bc.bc_rows, bc.bc_vals)
    wd = w0 + spla.spsolve(K.tocsc(), b)
    w = np.asarray(sol)
    err = np.linalg.norm(w - wd) / np.linalg.norm(wd)
    print("coupled gmres vs direct rel err", err, solver.stats)
    assert err < 1e-9
    N = prob.num_total_nodes
    assert u.shape == (N, 3) and T.shape == (N, 1)
    assert np.allclose(T.reshape(-1), wd[3 * N:], rtol=1e-9, atol=1e-9)
    assert np.allclose(u.reshape(-1), wd[:3 * N], rtol=1e-9, atol=1e-9)


def test_coupled_gradients_match_finite_differences():
    """d(w.w)/dE (cell variable) and d(w.w)/d(surface loads) through the
    adjoint on J_bc^T (the transpose of the non-symmetric coupled
    Jacobian) against central differences."""
    import torch
    import feax_amd as fe
    prob, bc, E0, tz0, qz0 = _setup(dims=(3, 2, 2))
    opts = fe.KrylovSolverOptions(solver="gmres", tol=1e-13, atol=0.0, restart=60, use_jacobi_preconditioner=True)
    solver = fe.create_solver(prob, bc, opts, linear=True, return_solution=False)

    def J(Ev, tv, qv):
        with torch.no_grad():
            w = solver(fe.TracedParams(volume_vars=(torch.as_tensor(Ev, device="cuda"),),
                                       surface_vars=[(torch.as_tensor(tv, device="cuda"),),
                                                     (torch.as_tensor(qv, device="cuda"),)]))
            return float((w * w).sum())

    print("case: coupled (u, T) gradients", flush=True)
    Et = torch.as_tensor(E0, device="cuda").requires_grad_(True)
    tt = torch.as_tensor(tz0, device="cuda").requires_grad_(True)
    qt = torch.as_tensor(qz0, device="cuda").requires_grad_(True)
    w = solver(fe.TracedParams(volume_vars=(Et,), surface_vars=[(tt,), (qt,)]))
    (w * w).sum().backward()
    gE, gt, gq = _np(Et.grad), _np(tt.grad), _np(qt.grad)
    rng = np.random.default_rng(11)
    for i in rng.choice(len(E0), 4, replace=False):
        Ep, Em = E0.copy(), E0.copy()
        Ep[i] += 1e-5
        Em[i] -= 1e-5
        fd = (J(Ep, tz0, qz0) - J(Em, tz0, qz0)) / 2e-5
        assert abs(fd - gE[i]) <= 1e-6 * np.abs(gE).max() + 1e-8 * abs(fd), (i, fd, gE[i])
    for arr, g, which in ((tz0, gt, 0), (qz0, gq, 1)):
        for f, q in zip(rng.choice(arr.shape[0], 2, replace=False), rng.choice(4, 2)):
            # w is affine in the surface loads, J quadratic: the central
            # difference is exact for any step, and a large one keeps the
            # solver's own error out of it
            ap, am = arr.copy(), arr.copy()
            ap[f, q] += 1e-2
            am[f, q] -= 1e-2
            args_p = (E0, ap, qz0) if which == 0 else (E0, tz0, ap)
            args_m = (E0, am, qz0) if which == 0 else (E0, tz0, am)
            fd = (J(*args_p) - J(*args_m)) / 2e-2
            assert abs(fd - g[f, q]) <= 1e-6 * np.abs(g).max() + 1e-8 * abs(fd), (which, f, q, fd, g[f, q])


def test_coupled_newton_matches_linear_solve_and_its_gradient():
    """linear=False on the (linear) coupled problem: one Newton step to the
    linear solution; d(w.w)/dE through the Newton adjoint (J_bc^T by GMRES)
    equals the linear path's."""
    import torch
    import feax_amd as fe
    prob, bc, E0, tz0, qz0 = _setup(dims=(3, 2, 2))
    opts = fe.KrylovSolverOptions(solver="gmres", tol=1e-13, atol=0.0, restart=60, use_jacobi_preconditioner=True)
    lin = fe.create_solver(prob, bc, opts, linear=True, return_solution=False)
    nwt = fe.create_solver(prob, bc, opts, newton_options=fe.NewtonOptions(tol=1e-10, rel_tol=1e-12), linear=False,
                           return_solution=False)
    print("case: coupled (u, T) Newton vs linear", flush=True)
    out = []
    for s in (lin, nwt):
        Et = torch.as_tensor(E0, device="cuda").requires_grad_(True)
        w = s(fe.TracedParams(volume_vars=(Et,), surface_vars=[(torch.as_tensor(tz0, device="cuda"),),
                                                                 (torch.as_tensor(qz0, device="cuda"),)]))
        (w * w).sum().backward()
        out.append((_np(w), _np(Et.grad)))
    (wl, gl), (wn, gn) = out
    assert np.abs(wn - wl).max() <= 1e-9 * np.abs(wl).max()
    assert np.abs(gn - gl).max() <= 1e-8 * np.abs(gl).max()
    with pytest.raises(NotImplementedError):
        fe.create_solver(prob, bc, fe.DirectSolverOptions(solver="cudss"), linear=True)(
            fe.TracedParams(volume_vars=(E0,), surface_vars=[(tz0,), (qz0,)]))
