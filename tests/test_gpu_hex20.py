"""GPU parity of the HEX20 kernels against the HEX20 oracle (oracle/hex20.py):
CSR assembly (FP64 / FP32, linear elastic node variable, SIMP cell
variable, Neo-Hookean) on a distorted, renumbered mesh; the residual with
dead traction and Dirichlet rows; linear solves (CG FP64, mixed precision,
direct) against scipy's direct solve; Neo-Hookean Newton load steps against
the oracle Newton; gradients (parameter and traction VJPs) against central
differences, and the fused compliance sensitivity against autograd."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EPS64 = np.finfo(np.float64).eps
EPS32 = np.finfo(np.float32).eps
NU = 0.3


def _np(t):
    return t.detach().cpu().numpy()


def _mesh(dims, distort=0.0, seed=0, shuffle=False):
    """Structured HEX20 box; interior corners moved by up to `distort`, edge
    nodes kept at edge midpoints; optionally nodes and cells renumbered."""
    import feax_amd as fe
    from oracle import hex20
    m = fe.hex20_mesh(dims)
    pts, cells = m.points.copy(), np.asarray(m.cells, np.int64)
    rng = np.random.default_rng(seed)
    if distort:
        inner = np.all((pts > 1e-9) & (pts < np.array(dims, float) - 1e-9), axis=1)
        corners = np.unique(cells[:, :8])
        sel = corners[inner[corners]]
        pts[sel] += rng.uniform(-distort, distort, (len(sel), 3))
        X = pts[cells]
        pts[cells[:, 8:].reshape(-1)] = (0.5 * (X[:, hex20.EDGES[:, 0]] + X[:, hex20.EDGES[:, 1]])).reshape(-1, 3)
    if shuffle:
        perm = rng.permutation(len(pts))
        pts, cells = pts[perm], np.argsort(perm)[cells][rng.permutation(len(cells))]
    return pts, cells


def _problem(kind, dims=(4, 2, 2), distort=0.12, seed=0, shuffle=True, load=True):
    import feax_amd as fe
    pts, cells = _mesh(dims, distort, seed, shuffle)
    L = float(dims[0])

    class P(fe.Problem):
        def get_tensor_map(self):
            if kind == "simp":
                return fe.materials.SIMPLinearElastic(nu=NU, E0=1.0, Emin=1e-3, penal=3.0)
            if kind == "neo_hookean":
                return None
            return fe.materials.IsotropicLinearElastic(nu=NU)

        def get_energy_density(self):
            return fe.materials.CompressibleNeoHookean(kappa=10.0) if kind == "neo_hookean" else None

        def get_surface_maps(self):
            return [fe.materials.SurfaceTraction((0.0, 0.0, 1.0))]

    locs = [lambda p: np.isclose(p[0], L)] if load else []
    prob = P(fe.Mesh(pts, cells, ele_type="HEX20"), vec=3, dim=3, ele_type="HEX20", location_fns=locs)
    bc = fe.DirichletBCConfig([fe.DirichletBCSpec(lambda p: np.isclose(p[0], 0.0), "all", 0.0)]).create_bc(prob)
    return prob, bc


def _oracle(prob):
    from oracle import hex20
    return hex20.Model(prob.points, prob.cells.astype(np.int64), prob.boundary_inds_list)


def _mat(kind):
    from oracle import hex8
    if kind == "simp":
        return hex8.Material("simp", nu=NU, E0=1.0, Emin=1e-3, penal=3.0)
    if kind == "neo_hookean":
        return hex8.Material("neo_hookean", kappa=10.0)
    return hex8.Material("linear_elastic", nu=NU)


def _param(kind, prob, seed=1):
    rng = np.random.default_rng(seed)
    if kind == "linear_elastic":
        return rng.uniform(50.0, 90.0, prob.num_total_nodes)      # node variable
    if kind == "simp":
        return rng.uniform(0.3, 1.0, prob.num_cells)
    return rng.uniform(0.5, 1.5, prob.num_cells)


@pytest.mark.parametrize("kind", ["linear_elastic", "simp", "neo_hookean"])
@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_hex20_assembly_matches_oracle(kind, dtype):
    import torch
    import feax_amd as fe
    from feax_amd.assembler import assemble_csr
    from oracle import hex8
    prob, bc = _problem(kind)
    m = _oracle(prob)
    ts = fe.TracedStructure.from_problem(prob)
    param = _param(kind, prob)
    u = 0.03 * np.random.default_rng(2).standard_normal(m.n) if kind == "neo_hookean" else np.zeros(m.n)
    dt = torch.float64 if dtype == "f64" else torch.float32
    print("case: HEX20 assembly", kind, dtype, flush=True)
    A = assemble_csr(prob, ts, fe.TracedParams(volume_vars=(param,)), torch.as_tensor(u, device="cuda"), bc, True, dt)
    assert np.array_equal(_np(A.indptr), m.csr["indptr"]) and np.array_equal(_np(A.indices), m.csr["indices"])
    mask = np.zeros(m.n, bool)
    mask[bc.bc_rows] = True
    ref = hex8.apply_bc_csr(m.csr_values(u, param, _mat(kind)), m.csr, mask, bc.bc_rows, True)
    K = m.scipy_csr(ref)
    scale = np.repeat(np.abs(K).max(axis=1).toarray().ravel(), np.diff(K.indptr))
    err = np.abs(_np(A.data).astype(np.float64) - ref) / scale
    tol = 64 * (EPS64 if dtype == "f64" else EPS32)
    print(kind, dtype, "assembly max err / row max", err.max())
    assert err.max() <= tol


def test_hex20_jacobi_and_difference_form():
    """dinv from the assembly pass and the FP32 difference-form operator
    (applied to FP64 vectors) against the oracle matrix."""
    import torch
    import feax_amd as fe
    from feax_amd.assembler import assemble_csr
    from oracle import hex8
    prob, bc = _problem("simp")
    m = _oracle(prob)
    ts = fe.TracedStructure.from_problem(prob)
    rho = _param("simp", prob)
    print("case: HEX20 difference form", flush=True)
    A, dinv = assemble_csr(prob, ts, fe.TracedParams(volume_vars=(rho,)), None, bc, True, torch.float32,
                           diff_form=True, jacobi_shift=1e-12)
    mask = np.zeros(m.n, bool)
    mask[bc.bc_rows] = True
    K = m.scipy_csr(hex8.apply_bc_csr(m.csr_values(np.zeros(m.n), rho, _mat("simp")), m.csr, mask, bc.bc_rows))
    assert np.allclose(_np(dinv), 1.0 / (K.diagonal() + 1e-12), rtol=1e-6)
    x = 1.0 + 1e-2 * np.random.default_rng(3).standard_normal(m.n)
    y = _np(A.matvec(torch.as_tensor(x, device="cuda")))
    ref = K @ x
    err = np.abs(y - ref).max() / (abs(K) @ np.abs(x - 1.0) + np.abs(ref)).max()
    print("diff3 operator rel err", err)
    assert err <= 64 * EPS32


@pytest.mark.parametrize("kind", ["linear_elastic", "neo_hookean"])
def test_hex20_residual_matches_oracle(kind):
    import torch
    import feax_amd as fe
    from feax_amd.assembler import residual
    from oracle import hex8
    prob, bc = _problem(kind)
    m = _oracle(prob)
    ts = fe.TracedStructure.from_problem(prob)
    param = _param(kind, prob)
    u = 0.03 * np.random.default_rng(4).standard_normal(m.n)
    t = fe.TracedParams.create_spatially_varying_surface_var(prob, lambda p: 0.3 + 0.1 * p[1] * p[2])
    tp = fe.TracedParams(volume_vars=(param,), surface_vars=[(t,)])
    print("case: HEX20 residual", kind, flush=True)
    r = _np(residual(prob, ts, tp, torch.as_tensor(u, device="cuda"), bc))
    ref = m.residual(u, param, _mat(kind), [np.asarray(t)[:, :, None] * np.array([0.0, 0.0, 1.0])])
    ref = hex8.apply_bc_res(ref, u, bc.bc_rows, bc.bc_vals)
    err = np.abs(r - ref).max() / np.abs(ref).max()
    print(kind, "residual rel err", err)
    assert err <= 1e-13


@pytest.mark.parametrize("opts", ["cg", "mixed", "direct"])
def test_hex20_linear_solve_matches_direct_solve(opts):
    import scipy.sparse.linalg as spla
    import feax_amd as fe
    from oracle import hex8, krylov
    prob, bc = _problem("simp", dims=(6, 2, 3))
    m = _oracle(prob)
    rho = _param("simp", prob)
    t = 0.5
    tp = fe.TracedParams(volume_vars=(rho,), surface_vars=[(fe.TracedParams.create_uniform_surface_var(prob, t),)])
    if opts == "direct":
        o = fe.DirectSolverOptions(solver="cudss")
    else:
        o = fe.KrylovSolverOptions(solver="cg", tol=1e-11, atol=0.0, use_jacobi_preconditioner=True,
                                   precision="mixed" if opts == "mixed" else "fp64")
    print("case: HEX20 linear solve", opts, flush=True)
    u = np.asarray(fe.create_solver(prob, bc, o, linear=True)(tp))
    _, info = krylov.linear_solve(m, _mat("simp"), rho, [hex8.constant_traction(prob.boundary_inds_list[0], t)],
                                  bc.bc_rows, bc.bc_vals, maxiter=1)
    ud = spla.spsolve(info["K"].tocsc(), info["b"])
    err = np.linalg.norm(u - ud) / np.linalg.norm(ud)
    print(opts, "hex20 solve rel err vs scipy", err)
    assert err < 1e-8


def test_hex20_newton_matches_oracle_newton():
    import torch
    import feax_amd as fe
    from oracle import newton as onewton
    dims = (2, 2, 3)
    prob, _ = _problem("neo_hookean", dims=dims, load=False, shuffle=False, distort=0.0)
    nz = dims[2]
    bc = fe.DirichletBCConfig([fe.DirichletBCSpec(lambda p: np.isclose(p[2], 0.0), "all", 0.0),
                               fe.DirichletBCSpec(lambda p: np.isclose(p[2], nz), "z", 0.0)]).create_bc(prob)
    node, comp = bc.bc_rows // 3, bc.bc_rows % 3
    is_top = np.isclose(prob.points[node, 2], nz) & (comp == 2)
    m = _oracle(prob)
    mu = _param("neo_hookean", prob)
    solver = fe.create_solver(prob, bc, fe.DirectSolverOptions(solver="cudss"),
                              newton_options=fe.NewtonOptions(tol=1e-10, rel_tol=1e-12), linear=False,
                              return_solution=False)
    tp = fe.TracedParams(volume_vars=(mu,))
    u = torch.zeros(m.n, dtype=torch.float64, device="cuda")
    uo = np.zeros(m.n)
    rows = torch.as_tensor(bc.bc_rows.astype(np.int64), device="cuda")
    for s in (1, 2):
        print("case: HEX20 Newton step", s, flush=True)
        vals = np.where(is_top, -0.05 * nz * s, 0.0)
        u[rows] = torch.as_tensor(vals, device="cuda")
        u = solver(tp, u, bc.replace_vals(vals))
        uo[bc.bc_rows] = vals
        uo, it_o, _ = onewton.newton(m, _mat("neo_hookean"), mu, bc.bc_rows, vals, uo, tol=1e-10, rel_tol=1e-12)
        err = np.linalg.norm(_np(u) - uo) / np.linalg.norm(uo)
        print("hex20 newton step", s, "iters", solver.stats.newton_iters, "oracle", it_o, "err", err)
        assert err < 1e-9 and solver.stats.newton_iters == it_o


@pytest.mark.parametrize("kind", ["simp", "linear_elastic"])
def test_hex20_gradients_match_finite_differences(kind):
    """d(u.u)/dp through the adjoint (HEX20 parameter VJP; cell rho for SIMP,
    node E for linear elasticity) and d(u.u)/dt through the traction VJP
    (4 values per face) against central differences of re-solved states;
    the fused compliance sensitivity against the autograd gradient."""
    import torch
    import feax_amd as fe
    prob, bc = _problem(kind, dims=(3, 2, 2))
    p0 = _param(kind, prob)
    opts = fe.KrylovSolverOptions(solver="cg", tol=1e-13, atol=0.0, use_jacobi_preconditioner=True)
    solver = fe.create_solver(prob, bc, opts, linear=True, return_solution=False)
    t0 = fe.TracedParams.create_uniform_surface_var(prob, 0.5)

    def J(pv, tv):
        with torch.no_grad():
            u = solver(fe.TracedParams(volume_vars=(torch.as_tensor(pv, device="cuda"),),
                                       surface_vars=[(torch.as_tensor(tv, device="cuda"),)]))
            return float((u * u).sum())

    print("case: HEX20 gradients", kind, flush=True)
    r = torch.as_tensor(p0, device="cuda").requires_grad_(True)
    tt = torch.as_tensor(t0, device="cuda").requires_grad_(True)
    u = solver(fe.TracedParams(volume_vars=(r,), surface_vars=[(tt,)]))
    (u * u).sum().backward()
    g, gt = _np(r.grad), _np(tt.grad)
    rng = np.random.default_rng(9)
    for i in rng.choice(len(p0), 5, replace=False):
        eps = 1e-5 * max(1.0, abs(p0[i]))
        pp, pm = p0.copy(), p0.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (J(pp, t0) - J(pm, t0)) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * np.abs(g).max() + 1e-9 * abs(fd), (i, fd, g[i])
    assert gt.shape == t0.shape
    for f, q in zip(rng.choice(t0.shape[0], 3, replace=False), rng.choice(4, 3)):
        tp_, tm_ = t0.copy(), t0.copy()
        tp_[f, q] += 1e-5
        tm_[f, q] -= 1e-5
        fd = (J(p0, tp_) - J(p0, tm_)) / 2e-5
        assert abs(fd - gt[f, q]) <= 1e-6 * np.abs(gt).max(), (f, q, fd, gt[f, q])
    if kind == "simp":
        from feax_amd.gene import compliance_and_sensitivity, create_compliance_fn
        tp = fe.TracedParams(volume_vars=(torch.as_tensor(p0, device="cuda"),), surface_vars=[(t0,)])
        with torch.no_grad():
            us = solver(tp)
        C, dC = compliance_and_sensitivity(prob, tp, us, surface_load_params=0.5)
        r2 = torch.as_tensor(p0, device="cuda").requires_grad_(True)
        u2 = solver(fe.TracedParams(volume_vars=(r2,), surface_vars=[(t0,)]))
        create_compliance_fn(prob, surface_load_params=0.5)(u2).backward()
        ref = _np(r2.grad)
        assert np.abs(_np(dC) - ref).max() <= 1e-8 * np.abs(ref).max()
