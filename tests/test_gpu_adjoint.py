"""GPU parity of the adjoint kernels and of gradients through the solver.

Kernels vs the oracle (oracle/adjoint.py, itself pinned by central
differences and the SIMP closed form); torch.autograd through create_solver
vs the fused sensitivity kernel and vs central differences of GPU solves.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy()


def _problem(kind, dims=(5, 3, 3), seed=11, distort=0.15, load=False):
    import feax_amd as fe
    from oracle import structure as st
    rng = np.random.default_rng(seed)
    pts, cells = st.structured_grid(dims, (1.0, 1.2, 0.8))
    pts0 = pts.copy()
    pts = pts + rng.uniform(-distort, distort, pts.shape)
    L = float(dims[0])

    laws = {"linear_elastic": lambda: fe.materials.IsotropicLinearElastic(nu=0.3),
            "simp": lambda: fe.materials.SIMPLinearElastic(nu=0.3, E0=1.0, Emin=1e-3, penal=3.0),
            "neo_hookean": lambda: fe.materials.CompressibleNeoHookean(kappa=10.0)}

    class P(fe.Problem):
        def get_tensor_map(self):
            return laws[kind]()

        def get_surface_maps(self):
            return [fe.materials.SurfaceTraction((0.3, -0.2, 1.0))] if load else []

    locs = [lambda p: np.isclose(p[0], L, atol=0.2)] if load else []
    # Faces are selected on the undistorted grid: keep x = L exactly planar.
    if load:
        pts[np.isclose(pts0[:, 0], L), 0] = L
    pts[np.isclose(pts0[:, 0], 0.0), 0] = 0.0
    prob = P(fe.Mesh(pts, cells), vec=3, dim=3, location_fns=locs)
    bc = fe.DirichletBCConfig([fe.DirichletBCSpec(lambda p: np.isclose(p[0], 0.0), "all", 0.0)]).create_bc(prob)
    return prob, bc, rng


def _oracle_mat(kind):
    from oracle import hex8
    return hex8.Material(kind, nu=0.3, E0=1.0, Emin=1e-3, penal=3.0, kappa=10.0)


@pytest.mark.parametrize("kind,loc", [("linear_elastic", "node"), ("linear_elastic", "cell"), ("simp", "cell"),
                                      ("simp", "node"), ("neo_hookean", "cell")])
def test_param_vjp_matches_oracle(kind, loc):
    import torch
    import feax_amd as fe
    from feax_amd.assembler import param_vjp
    from oracle import adjoint, hex8
    prob, bc, rng = _problem(kind)
    size = prob.num_cells if loc == "cell" else prob.num_total_nodes
    p = rng.uniform(0.4, 1.0, size)
    u = (0.05 if kind == "neo_hookean" else 1.0) * rng.standard_normal(prob.num_total_dofs_all_vars)
    lam = rng.standard_normal(u.size)
    tp = fe.TracedParams(volume_vars=(p,))
    ts = fe.TracedStructure.from_problem(prob)
    g = _np(param_vjp(prob, ts, tp, torch.as_tensor(u, device="cuda"), torch.as_tensor(lam, device="cuda"), bc,
                      scale=-1.0))
    m = hex8.Model(prob.points, prob.cells.astype(np.int64), structure=False)
    ref = -adjoint.param_vjp(m, u, lam, p, _oracle_mat(kind), bc_mask=bc.bc_mask)
    err = np.abs(g - ref).max() / np.abs(ref).max()
    print(kind, loc, "max rel err", err)
    assert err < 1e-12


def test_compliance_sensitivity_kernel_matches_closed_form():
    import torch
    import feax_amd as fe
    from feax_amd.assembler import compliance_sensitivity
    from oracle import adjoint, hex8
    prob, bc, rng = _problem("simp", dims=(6, 4, 3), seed=12)
    rho = rng.uniform(0.3, 1.0, prob.num_cells)
    u = rng.standard_normal(prob.num_total_dofs_all_vars)
    ts = fe.TracedStructure.from_problem(prob)
    dc = _np(compliance_sensitivity(prob, ts, fe.TracedParams(volume_vars=(rho,)), torch.as_tensor(u, device="cuda")))
    m = hex8.Model(prob.points, prob.cells.astype(np.int64), structure=False)
    ref = adjoint.compliance_sens_closed_form(m, u, rho, _oracle_mat("simp"))
    assert np.abs(dc - ref).max() / np.abs(ref).max() < 1e-12


def test_traction_vjp_matches_oracle():
    import torch
    import feax_amd as fe
    from feax_amd.assembler import traction_vjp
    from oracle import adjoint, hex8
    prob, bc, rng = _problem("linear_elastic", load=True, seed=13)
    assert len(prob.boundary_inds_list[0]) > 0
    lam = rng.standard_normal(prob.num_total_dofs_all_vars)
    ts = fe.TracedStructure.from_problem(prob)
    g = _np(traction_vjp(prob, ts, torch.as_tensor(lam, device="cuda"), bc, scale=1.0)[0])
    m = hex8.Model(prob.points, prob.cells.astype(np.int64), structure=False)
    ref = adjoint.traction_vjp(m, prob.boundary_inds_list[0], (0.3, -0.2, 1.0), lam, bc_mask=bc.bc_mask)
    assert np.abs(g - ref).max() / np.abs(ref).max() < 1e-12


def test_strain_energy_matches_oracle():
    import torch
    import feax_amd as fe
    from feax_amd.assembler import strain_energy
    from oracle import adjoint, hex8
    prob, bc, rng = _problem("neo_hookean", seed=14)
    mu = rng.uniform(0.5, 1.5, prob.num_cells)
    u = 0.05 * rng.standard_normal(prob.num_total_dofs_all_vars)
    ts = fe.TracedStructure.from_problem(prob)
    Pi, dmu = strain_energy(prob, ts, fe.TracedParams(volume_vars=(mu,)), torch.as_tensor(u, device="cuda"),
                            with_partial=True)
    m = hex8.Model(prob.points, prob.cells.astype(np.int64), structure=False)
    Pr, dr = adjoint.neo_hookean_energy_total(m, u, mu, 10.0)
    assert abs(float(Pi) - Pr) <= 1e-12 * abs(Pr)
    assert np.abs(_np(dmu) - dr).max() <= 1e-12 * np.abs(dr).max()


def _simp_cantilever(dims=(12, 4, 4)):
    """SIMP cantilever: x = 0 clamped, unit load on a patch at mid right face."""
    import feax_amd as fe
    nx, ny, nz = dims
    yc, zc = ny / 2.0, nz / 2.0

    class Simp(fe.Problem):
        def get_tensor_map(self):
            return fe.materials.SIMPLinearElastic(nu=0.3, E0=1.0, Emin=1e-9, penal=3.0)

        def get_surface_maps(self):
            return [fe.materials.SurfaceTraction((0.0, 0.0, 1.0))]

    mesh = fe.StructuredGrid(dims, spacing=(1.0, 1.0, 1.0)).to_mesh()
    patch = [lambda p: np.isclose(p[0], nx) & (np.abs(p[1] - yc) <= 1.0 + 1e-9) & (np.abs(p[2] - zc) <= 1.0 + 1e-9)]
    prob = Simp(mesh, vec=3, dim=3, location_fns=patch)
    bc = fe.DirichletBCConfig([fe.DirichletBCSpec(lambda p: np.isclose(p[0], 0.0), "all", 0.0)]).create_bc(prob)
    return prob, bc


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_autograd_compliance_gradient(precision):
    """rho.grad from C = compliance_fn(solver(rho)) vs the fused self-adjoint
    kernel, the oracle closed form on a direct solve, and central differences
    of GPU solves on 4 seeded elements."""
    import torch
    import scipy.sparse.linalg as spla
    import feax_amd as fe
    from feax_amd.gene import compliance_and_sensitivity, create_compliance_fn
    from oracle import adjoint, hex8
    prob, bc = _simp_cantilever()
    rng = np.random.default_rng(20261203)
    rho0 = rng.uniform(0.3, 1.0, prob.num_cells)
    nf = len(prob.boundary_inds_list[0])
    t = fe.TracedParams.create_uniform_surface_var(prob, 0.25)
    assert nf == 4
    opts = fe.KrylovSolverOptions(solver="cg", tol=1e-11, atol=0.0, maxiter=20000, use_jacobi_preconditioner=True,
                                  precision=precision)
    solver = fe.create_solver(prob, bc, opts, linear=True, return_solution=False)
    comp = create_compliance_fn(prob, surface_load_params=0.25)
    rho = torch.tensor(rho0, device="cuda", requires_grad=True)
    u = solver(fe.TracedParams(volume_vars=(rho,), surface_vars=[(t,)]))
    C = comp(u)
    C.backward()
    g = _np(rho.grad)
    # fused self-adjoint kernel on the same state
    C2, dc = compliance_and_sensitivity(prob, fe.TracedParams(volume_vars=(rho0,)), u.detach(), 0.25)
    assert abs(float(C2) - float(C)) <= 1e-14 * abs(float(C))
    np.testing.assert_allclose(g, _np(dc), rtol=1e-6, atol=1e-9 * np.abs(g).max())
    # oracle: direct solve + closed form
    m = hex8.Model(prob.points, prob.cells.astype(np.int64), prob.boundary_inds_list)
    mat = hex8.Material("simp", nu=0.3, E0=1.0, Emin=1e-9, penal=3.0)
    K = m.scipy_csr(hex8.apply_bc_csr(m.csr_values(np.zeros(m.n), rho0, mat), m.csr, bc.bc_mask, bc.bc_rows))
    tr = [hex8.constant_traction(prob.boundary_inds_list[0], 0.25)]
    f = -m.residual(np.zeros(m.n), rho0, mat, tr)
    f[bc.bc_rows] = 0.0
    ustar = spla.spsolve(K.tocsc(), f)
    ref = adjoint.compliance_sens_closed_form(m, ustar, rho0, mat)
    err = np.abs(g - ref).max() / np.abs(ref).max()
    print(precision, "grad vs oracle closed form", err)
    assert err < 1e-6
    # central differences of the GPU solve
    idx = rng.choice(prob.num_cells, 4, replace=False)
    eps = 1e-4

    def Cof(r):
        with torch.no_grad():
            return float(comp(solver(fe.TracedParams(volume_vars=(torch.as_tensor(r, device="cuda"),),
                                                     surface_vars=[(t,)]))))
    for i in idx:
        rp, rm = rho0.copy(), rho0.copy()
        rp[i] += eps
        rm[i] -= eps
        fd = (Cof(rp) - Cof(rm)) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-5 * abs(g[i]) + 1e-9 * np.abs(g).max(), (i, fd, g[i])


def test_autograd_general_objective_node_modulus_and_traction():
    """L = w . u with a random w: dL/dE (node var) and dL/dt (surface var) vs
    central differences of GPU solves (u is linear in t: exact differences)."""
    import torch
    import feax_amd as fe
    from conftest import cantilever
    prob, bc, tp0 = cantilever((8, 3, 3), (1.0, 1.0, 1.0), 8.0)
    rng = np.random.default_rng(21)
    E0 = rng.uniform(50e3, 90e3, prob.num_total_nodes)
    t0 = rng.uniform(0.5e-3, 1.5e-3, (len(prob.boundary_inds_list[0]), 4))
    w = torch.as_tensor(rng.standard_normal(prob.num_total_dofs_all_vars), device="cuda")
    opts = fe.KrylovSolverOptions(solver="cg", tol=1e-12, atol=0.0, maxiter=20000, use_jacobi_preconditioner=True)
    solver = fe.create_solver(prob, bc, opts, linear=True, return_solution=False)
    E = torch.tensor(E0, device="cuda", requires_grad=True)
    t = torch.tensor(t0, device="cuda", requires_grad=True)
    L = torch.dot(w, solver(fe.TracedParams(volume_vars=(E,), surface_vars=[(t,)])))
    L.backward()
    gE, gt = _np(E.grad), _np(t.grad)

    def Lof(Ev, tv):
        with torch.no_grad():
            return float(torch.dot(w, solver(fe.TracedParams(volume_vars=(torch.as_tensor(Ev, device="cuda"),),
                                                             surface_vars=[(torch.as_tensor(tv, device="cuda"),)]))))
    for i in rng.choice(prob.num_total_nodes, 3, replace=False):
        h = 1e-3 * E0[i]
        Ep, Em = E0.copy(), E0.copy()
        Ep[i] += h
        Em[i] -= h
        fd = (Lof(Ep, t0) - Lof(Em, t0)) / (2 * h)
        assert abs(fd - gE[i]) <= 1e-5 * np.abs(gE).max(), (i, fd, gE[i])
    for k in rng.choice(t0.size, 3, replace=False):
        tp_, tm_ = t0.copy().reshape(-1), t0.copy().reshape(-1)
        tp_[k] += 1e-4
        tm_[k] -= 1e-4
        fd = (Lof(E0, tp_.reshape(t0.shape)) - Lof(E0, tm_.reshape(t0.shape))) / 2e-4
        assert abs(fd - gt.reshape(-1)[k]) <= 1e-6 * np.abs(gt).max(), (k, fd, gt.reshape(-1)[k])
