"""GPU parity of the HEX8 thermal strain (thermo-elastic laws) against the
oracle (oracle/thermoelastic.py): the residual with a node temperature
field, free expansion on roller supports (exact u = alpha dT x), gradients
with respect to the modulus and the temperature against central
differences, and the staggered thermal -> mechanical chain of feax's
examples/basic/thermo_mech_coupling.py with the gradient flowing back into
the thermal conductivity."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy()


def _mesh(dims=(6, 3, 3), seed=0, distort=True):
    import feax_amd as fe
    from oracle import structure as st
    pts, cells = st.structured_grid(dims, (1.0, 1.0, 1.0))
    L = np.array(dims, float)
    if distort:
        inner = np.all((pts > 1e-9) & (pts < L - 1e-9), axis=1)
        pts = pts.copy()
        pts[inner] += np.random.default_rng(seed).uniform(-0.12, 0.12, (inner.sum(), 3))
    return fe.Mesh(pts, cells), L


def _mech(mesh, alpha=2e-3, simp=False):
    import feax_amd as fe

    class TE(fe.Problem):
        def get_tensor_map(self):
            if simp:
                return fe.materials.SIMPThermoElastic(nu=0.3, E0=100.0, Emin=1e-3, penal=3.0, alpha=alpha)
            return fe.materials.ThermoElastic(nu=0.3, alpha=alpha, t_ref=10.0)

    return TE(mesh, vec=3, dim=3)


def test_thermoelastic_residual_matches_oracle():
    import torch
    import feax_amd as fe
    from feax_amd.assembler import residual
    from oracle import hex8, thermoelastic
    mesh, _ = _mesh()
    prob = _mech(mesh)
    bc = fe.DirichletBCConfig([fe.DirichletBCSpec(lambda p: np.isclose(p[0], 0.0), "all", 0.0)]).create_bc(prob)
    m = hex8.Model(prob.points, prob.cells.astype(np.int64))
    rng = np.random.default_rng(3)
    E = rng.uniform(50.0, 90.0, prob.num_cells)
    T = rng.uniform(0.0, 100.0, prob.num_total_nodes)
    u = 0.01 * rng.standard_normal(m.n)
    ts = fe.TracedStructure.from_problem(prob)
    print("thermoelastic residual", flush=True)
    r = _np(residual(prob, ts, fe.TracedParams(volume_vars=(E, T)), torch.as_tensor(u, device="cuda"), bc))
    ref = thermoelastic.residual(m, u, E, hex8.Material("linear_elastic", nu=0.3), T, 2e-3, 10.0)
    ref = hex8.apply_bc_res(ref, u, bc.bc_rows, bc.bc_vals)
    err = np.abs(r - ref).max() / np.abs(ref).max()
    print("residual rel err", err)
    assert err <= 1e-13


@pytest.mark.parametrize("opts", ["cg", "direct"])
def test_free_expansion_on_rollers_is_exact(opts):
    import feax_amd as fe
    mesh, L = _mesh(distort=True)
    prob = _mech(mesh, alpha=1e-3)
    bc = fe.DirichletBCConfig([fe.DirichletBCSpec(lambda p: np.isclose(p[0], 0.0), "x", 0.0),
                               fe.DirichletBCSpec(lambda p: np.isclose(p[1], 0.0), "y", 0.0),
                               fe.DirichletBCSpec(lambda p: np.isclose(p[2], 0.0), "z", 0.0)]).create_bc(prob)
    o = fe.DirectSolverOptions(solver="cudss") if opts == "direct" else \
        fe.KrylovSolverOptions(solver="cg", tol=1e-12, atol=0.0, use_jacobi_preconditioner=True)
    dT = 60.0
    tp = fe.TracedParams(volume_vars=(fe.TracedParams.create_cell_var(prob, 70.0),
                                      np.full(prob.num_total_nodes, 10.0 + dT)))
    print("free expansion", opts, flush=True)
    u = np.asarray(fe.create_solver(prob, bc, o, linear=True)(tp)).reshape(-1, 3)
    exact = 1e-3 * dT * prob.points
    err = np.abs(u - exact).max() / np.abs(exact).max()
    print(opts, "free expansion err", err)
    assert err <= 1e-9


def test_thermoelastic_gradients_match_finite_differences():
    """d(u.u)/dE (cell) and d(u.u)/dT (node) through the adjoint (the
    modulus pullback with its thermal term, fx_hex8_thermal_vjp_f64)."""
    import torch
    import feax_amd as fe
    mesh, _ = _mesh((4, 2, 2))
    prob = _mech(mesh)
    bc = fe.DirichletBCConfig([fe.DirichletBCSpec(lambda p: np.isclose(p[0], 0.0), "all", 0.0)]).create_bc(prob)
    rng = np.random.default_rng(4)
    E0 = rng.uniform(50.0, 90.0, prob.num_cells)
    T0 = rng.uniform(0.0, 100.0, prob.num_total_nodes)
    solver = fe.create_solver(prob, bc, fe.KrylovSolverOptions(solver="cg", tol=1e-13, atol=0.0,
                              use_jacobi_preconditioner=True), linear=True, return_solution=False)

    def J(Ev, Tv):
        with torch.no_grad():
            u = solver(fe.TracedParams(volume_vars=(torch.as_tensor(Ev, device="cuda"),
                                                    torch.as_tensor(Tv, device="cuda"))))
            return float((u * u).sum())
    Et = torch.tensor(E0, device="cuda", requires_grad=True)
    Tt = torch.tensor(T0, device="cuda", requires_grad=True)
    print("thermoelastic gradients", flush=True)
    u = solver(fe.TracedParams(volume_vars=(Et, Tt)))
    (u * u).sum().backward()
    gE, gT = _np(Et.grad), _np(Tt.grad)
    for i in rng.choice(len(E0), 4, replace=False):
        p, m = E0.copy(), E0.copy()
        p[i] += 1e-4
        m[i] -= 1e-4
        fd = (J(p, T0) - J(m, T0)) / 2e-4
        assert abs(fd - gE[i]) <= 1e-6 * np.abs(gE).max(), (i, fd, gE[i])
    for i in rng.choice(len(T0), 4, replace=False):
        p, m = T0.copy(), T0.copy()
        p[i] += 1e-3
        m[i] -= 1e-3
        fd = (J(E0, p) - J(E0, m)) / 2e-3
        assert abs(fd - gT[i]) <= 1e-6 * np.abs(gT).max(), (i, fd, gT[i])


def test_thermal_to_mechanical_chain_gradient():
    """Conduction (hot and cold end faces, nodal conductivity) -> node
    temperature -> thermo-elastic SIMP solve on the same mesh; d(u.u)/dk
    through both solves against central differences."""
    import torch
    import feax_amd as fe
    mesh, L = _mesh((4, 2, 2), distort=False)

    class Thermal(fe.Problem):
        def get_tensor_map(self):
            return fe.materials.IsotropicConduction()

    th = Thermal(mesh, vec=1, dim=3)
    bc_th = fe.DirichletBCConfig([fe.DirichletBCSpec(lambda p: np.isclose(p[0], 0.0), "all", 100.0),
                                  fe.DirichletBCSpec(lambda p: np.isclose(p[0], L[0]), "all", 0.0)]).create_bc(th)
    me = _mech(mesh, alpha=1e-4, simp=True)
    bc_me = fe.DirichletBCConfig([fe.DirichletBCSpec(lambda p: np.isclose(p[0], 0.0), "all", 0.0)]).create_bc(me)
    opts = fe.KrylovSolverOptions(solver="cg", tol=1e-13, atol=0.0, use_jacobi_preconditioner=True)
    sT = fe.create_solver(th, bc_th, opts, linear=True, return_solution=False)
    sU = fe.create_solver(me, bc_me, opts, linear=True, return_solution=False)
    rng = np.random.default_rng(5)
    k0 = rng.uniform(0.5, 1.5, th.num_total_nodes)
    rho = rng.uniform(0.4, 1.0, me.num_cells)

    def chain(k):
        T = sT(fe.TracedParams(volume_vars=(k,)), fe.zero_like_initial_guess(th, bc_th))
        Tn = fe.TracedParams.node_var_from_solution(th, T)
        u = sU(fe.TracedParams(volume_vars=(torch.as_tensor(rho, device="cuda"), Tn)))
        return (u * u).sum()
    k = torch.tensor(k0, device="cuda", requires_grad=True)
    print("thermal -> mechanical chain", flush=True)
    chain(k).backward()
    g = _np(k.grad)
    assert np.abs(g).max() > 0
    for i in rng.choice(len(k0), 4, replace=False):
        p, m = k0.copy(), k0.copy()
        p[i] += 1e-5
        m[i] -= 1e-5
        with torch.no_grad():
            fd = (float(chain(torch.as_tensor(p, device="cuda"))) - float(chain(torch.as_tensor(m, device="cuda")))) / 2e-5
        assert abs(fd - g[i]) <= 1e-6 * np.abs(g).max(), (i, fd, g[i])
