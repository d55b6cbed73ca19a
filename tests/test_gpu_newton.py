"""GPU parity of the Neo-Hookean kernels and the Newton solver.

Tangent and residual vs the oracle (whose tangent is pinned by central
differences of its residual); Newton solutions vs the oracle Newton with
direct solves; dPi/dmu through torch.autograd vs the equilibrium identity
(partial derivative) and central differences of re-solved GPU states.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EPS64 = np.finfo(np.float64).eps


def _np(t):
    return t.detach().cpu().numpy()


def _nh_problem(dims, distort=0.0, seed=0):
    import feax_amd as fe
    from oracle import structure as st
    nz = dims[2]
    pts, cells = st.structured_grid(dims)
    if distort:
        rng = np.random.default_rng(seed)
        inner = ~(np.isclose(pts[:, 2], 0.0) | np.isclose(pts[:, 2], nz))
        pts = pts + rng.uniform(-distort, distort, pts.shape) * inner[:, None]

    class NH(fe.Problem):
        def get_energy_density(self):
            return fe.materials.CompressibleNeoHookean(kappa=10.0)

    prob = NH(fe.Mesh(pts, cells), vec=3, dim=3)
    bc = fe.DirichletBCConfig([fe.DirichletBCSpec(lambda p: np.isclose(p[2], 0.0), "all", 0.0),
                               fe.DirichletBCSpec(lambda p: np.isclose(p[2], nz), "z", 0.0)]).create_bc(prob)
    node, comp = bc.bc_rows // 3, bc.bc_rows % 3
    is_top = np.isclose(prob.points[node, 2], nz) & (comp == 2)
    return prob, bc, is_top


def _row_scale(K):
    return np.repeat(np.abs(K).max(axis=1).toarray().ravel(), np.diff(K.indptr))


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_nh_tangent_matches_oracle(dtype):
    import torch
    import feax_amd as fe
    from feax_amd.assembler import assemble_csr
    from oracle import hex8
    prob, bc, _ = _nh_problem((4, 3, 3), distort=0.12, seed=3)
    rng = np.random.default_rng(4)
    mu = rng.uniform(0.5, 1.5, prob.num_cells)
    u = 0.05 * rng.standard_normal(prob.num_total_dofs_all_vars)
    ts = fe.TracedStructure.from_problem(prob)
    tdt = torch.float64 if dtype == "f64" else torch.float32
    A = assemble_csr(prob, ts, fe.TracedParams(volume_vars=(mu,)), torch.as_tensor(u, device="cuda"), None, True, tdt)
    m = hex8.Model(prob.points, prob.cells.astype(np.int64))
    ref = m.csr_values(u, mu, hex8.Material("neo_hookean", kappa=10.0))
    K = m.scipy_csr(ref)
    err = np.abs(_np(A.data).astype(np.float64) - ref) / _row_scale(K)
    tol = 64 * EPS64 if dtype == "f64" else 2.0 ** -23
    print(dtype, "NH tangent max err / row max", err.max())
    assert err.max() <= tol


def test_nh_residual_matches_oracle():
    import torch
    import feax_amd as fe
    from feax_amd.assembler import residual
    from oracle import hex8
    prob, bc, _ = _nh_problem((4, 3, 3), distort=0.12, seed=5)
    rng = np.random.default_rng(6)
    mu = rng.uniform(0.5, 1.5, prob.num_cells)
    u = 0.05 * rng.standard_normal(prob.num_total_dofs_all_vars)
    ts = fe.TracedStructure.from_problem(prob)
    r = _np(residual(prob, ts, fe.TracedParams(volume_vars=(mu,)), torch.as_tensor(u, device="cuda")))
    m = hex8.Model(prob.points, prob.cells.astype(np.int64), structure=False)
    ref = m.residual(u, mu, hex8.Material("neo_hookean", kappa=10.0))
    assert np.abs(r - ref).max() <= 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize("precision", ["fp64", "mixed", "direct"])
def test_newton_load_steps_match_oracle(precision):
    """-10% z-compression of a block in 3 displacement steps (bottom clamped,
    top u_z prescribed, initial guess = previous state with the BC values)."""
    import torch
    import feax_amd as fe
    from oracle import hex8, newton as onewton
    dims = (4, 3, 5)
    prob, bc, is_top = _nh_problem(dims)
    rng = np.random.default_rng(20261203)
    mu = rng.uniform(0.5, 1.5, prob.num_cells)
    if precision == "direct":   # the oracle's own linear solver: a direct solve
        lin = fe.DirectSolverOptions(solver="cudss")
    else:
        lin = fe.KrylovSolverOptions(solver="cg", tol=1e-10, atol=0.0, maxiter=20000,
                                     use_jacobi_preconditioner=True, precision=precision)
    nopts = fe.NewtonOptions(tol=1e-10, rel_tol=1e-12)
    solver = fe.create_solver(prob, bc, lin, newton_options=nopts, linear=False, return_solution=False)
    m = hex8.Model(prob.points, prob.cells.astype(np.int64))
    mat = hex8.Material("neo_hookean", kappa=10.0)
    tp = fe.TracedParams(volume_vars=(mu,))
    u = torch.zeros(prob.num_total_dofs_all_vars, dtype=torch.float64, device="cuda")
    uo = np.zeros(m.n)
    rows = torch.as_tensor(bc.bc_rows.astype(np.int64), device="cuda")
    for s in range(1, 4):
        vals = np.where(is_top, -0.1 * dims[2] * s / 3, 0.0)
        bcs = bc.replace_vals(vals)
        u[rows] = torch.as_tensor(vals, device="cuda")
        u = solver(tp, u, bcs)
        uo[bc.bc_rows] = vals
        uo, it_o, _ = onewton.newton(m, mat, mu, bc.bc_rows, vals, uo, tol=1e-10, rel_tol=1e-12)
        st = solver.stats
        err = np.linalg.norm(_np(u) - uo) / np.linalg.norm(uo)
        print(precision, "step", s, "newton iters", st.newton_iters, "oracle", it_o, "res", st.res_history[-1],
              "err", err)
        assert err < 1e-9
        assert st.newton_iters <= it_o + (3 if precision == "mixed" else 1)


@pytest.mark.parametrize("linear", ["cg", "direct"])
def test_energy_gradient_through_newton(linear):
    """dPi/dmu of the displacement-controlled equilibrium via torch.autograd
    (adjoint through the Newton solver + energy VJP) == the partial dPi/dmu_e
    (R_free = 0, du_bc/dmu = 0), and == central differences of GPU re-solves."""
    import torch
    import feax_amd as fe
    from feax_amd.assembler import create_energy_fn, strain_energy
    dims = (3, 3, 4)
    prob, bc, is_top = _nh_problem(dims)
    rng = np.random.default_rng(7)
    mu0 = rng.uniform(0.5, 1.5, prob.num_cells)
    vals = np.where(is_top, -0.1 * dims[2], 0.0)
    bcs = bc.replace_vals(vals)
    lin = fe.DirectSolverOptions(solver="cudss") if linear == "direct" else \
        fe.KrylovSolverOptions(solver="cg", tol=1e-12, atol=0.0, maxiter=20000, use_jacobi_preconditioner=True)
    nopts = fe.NewtonOptions(tol=1e-11, rel_tol=1e-13)
    solver = fe.create_solver(prob, bcs, lin, newton_options=nopts, linear=False, return_solution=False)
    energy = create_energy_fn(prob)
    ig = torch.zeros(prob.num_total_dofs_all_vars, dtype=torch.float64, device="cuda")
    ig[torch.as_tensor(bc.bc_rows.astype(np.int64), device="cuda")] = torch.as_tensor(vals, device="cuda")
    mu = torch.tensor(mu0, device="cuda", requires_grad=True)
    tp = fe.TracedParams(volume_vars=(mu,))
    u = solver(tp, ig)
    Pi = energy(u, tp)
    Pi.backward()
    g = _np(mu.grad)
    ts = prob._default_ts
    _, partial = strain_energy(prob, ts, fe.TracedParams(volume_vars=(mu0,)), u.detach(), with_partial=True)
    partial = _np(partial)
    assert np.abs(g - partial).max() <= 1e-8 * np.abs(partial).max()
    eps = 1e-5

    def Pof(mv):
        with torch.no_grad():
            tpv = fe.TracedParams(volume_vars=(torch.as_tensor(mv, device="cuda"),))
            return float(strain_energy(prob, ts, tpv, solver(tpv, u.detach())))
    for e in rng.choice(prob.num_cells, 4, replace=False):
        mp, mm = mu0.copy(), mu0.copy()
        mp[e] += eps
        mm[e] -= eps
        fd = (Pof(mp) - Pof(mm)) / (2 * eps)
        assert abs(fd - g[e]) <= 1e-6 * np.abs(g).max(), (e, fd, g[e])


@pytest.mark.parametrize("kind", ["simp", "neo_hookean"])
def test_matfree_operator_matches_assembled_csr(kind):
    """fx_hex8_jvp_f64 (feax's matrix-free Krylov operator J_bc v) == the
    assembled BC-eliminated CSR times v (feax/assembler.py:1753-1754)."""
    import torch
    import feax_amd as fe
    from feax_amd.assembler import assemble_csr, jvp
    prob, bc, _ = _nh_problem((4, 3, 3), distort=0.12, seed=8)
    if kind == "simp":
        class S(fe.Problem):
            def get_tensor_map(self):
                return fe.materials.SIMPLinearElastic(nu=0.3, E0=1.0, Emin=1e-3, penal=3.0)
        prob = S(fe.Mesh(prob.points, prob.cells), vec=3, dim=3)
    rng = np.random.default_rng(9)
    p = rng.uniform(0.5, 1.5, prob.num_cells)
    u = 0.05 * rng.standard_normal(prob.num_total_dofs_all_vars)
    v = rng.standard_normal(prob.num_total_dofs_all_vars)
    ts = fe.TracedStructure.from_problem(prob)
    tp = fe.TracedParams(volume_vars=(p,))
    ud, vd = torch.as_tensor(u, device="cuda"), torch.as_tensor(v, device="cuda")
    A = assemble_csr(prob, ts, tp, ud, bc, True, torch.float64)
    ref = _np(A.matvec(vd))
    got = _np(jvp(prob, ts, tp, ud if kind == "neo_hookean" else None, vd, bc))
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()
