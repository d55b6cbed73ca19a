"""GPU parity of the scalar HEX8 conduction kernels (vec = 1) against the
conduction oracle (oracle/conduction.py): CSR assembly (FP64 / FP32; node
and cell conductivity, SIMP law), the residual with a per-point surface flux
and Dirichlet rows, linear solves (CG, direct) against scipy and the exact
bar profile, gradients (conductivity, flux) against central differences,
the fused thermal compliance sensitivity against autograd, and the
thermal-to-next-solve chaining of feax's tests/test_params_bridge.py."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EPS64 = np.finfo(np.float64).eps
EPS32 = np.finfo(np.float32).eps


def _np(t):
    return t.detach().cpu().numpy()


def _problem(kind, dims=(8, 4, 3), seed=0, flux=True):
    import feax_amd as fe
    from oracle import structure as st
    pts, cells = st.structured_grid(dims, (1.0, 1.0, 1.0))
    rng = np.random.default_rng(seed)
    L = np.array(dims, float)
    inner = np.all((pts > 1e-9) & (pts < L - 1e-9), axis=1)
    pts = pts.copy()
    pts[inner] += rng.uniform(-0.15, 0.15, (inner.sum(), 3))
    perm = rng.permutation(len(pts))
    pts, cells = pts[perm], np.argsort(perm)[cells][rng.permutation(len(cells))]

    class Thermal(fe.Problem):
        def get_tensor_map(self):
            if kind == "simp":
                return fe.materials.SIMPConduction(k0=1.0, kmin=1e-3, penal=3.0)
            return fe.materials.IsotropicConduction()

        def get_surface_maps(self):
            return [fe.materials.SurfaceFlux()]

    locs = [lambda p: np.isclose(p[0], L[0])] if flux else []
    prob = Thermal(fe.Mesh(pts, cells), vec=1, dim=3, location_fns=locs)
    bc = fe.DirichletBCConfig([fe.DirichletBCSpec(lambda p: np.isclose(p[0], 0.0), "all", 0.0)]).create_bc(prob)
    return prob, bc, L


def _oracle(prob):
    from oracle import conduction
    return conduction.Model(prob.points, prob.cells.astype(np.int64), prob.boundary_inds_list)


def _param(kind, prob, loc, seed=1):
    rng = np.random.default_rng(seed)
    size = prob.num_total_nodes if loc == "node" else prob.num_cells
    return rng.uniform(0.3, 1.0, size) if kind == "simp" else rng.uniform(0.5, 2.0, size)


def _okind(kind):
    return "simp" if kind == "simp" else "conduction"


@pytest.mark.parametrize("kind,loc", [("conduction", "cell"), ("conduction", "node"), ("simp", "cell")])
@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_conduction_assembly_matches_oracle(kind, loc, dtype):
    import torch
    import feax_amd as fe
    from feax_amd.assembler import assemble_csr
    from oracle import hex8
    prob, bc, _ = _problem(kind)
    m = _oracle(prob)
    ts = fe.TracedStructure.from_problem(prob)
    param = _param(kind, prob, loc)
    dt = torch.float64 if dtype == "f64" else torch.float32
    print("conduction assembly", kind, loc, dtype, flush=True)
    A, dinv = assemble_csr(prob, ts, fe.TracedParams(volume_vars=(param,)), None, bc, True, dt, jacobi_shift=1e-12)
    assert np.array_equal(_np(A.indptr), m.csr["indptr"]) and np.array_equal(_np(A.indices), m.csr["indices"])
    mask = np.zeros(m.n, bool)
    mask[bc.bc_rows] = True
    ref = hex8.apply_bc_csr(m.csr_values(param, _okind(kind)), m.csr, mask, bc.bc_rows, True)
    K = m.scipy_csr(ref)
    scale = np.repeat(np.abs(K).max(axis=1).toarray().ravel(), np.diff(K.indptr))
    err = np.abs(_np(A.data).astype(np.float64) - ref) / scale
    print(kind, loc, dtype, "assembly max err / row max", err.max())
    assert err.max() <= 64 * (EPS64 if dtype == "f64" else EPS32)
    assert np.allclose(_np(dinv), 1.0 / (K.diagonal() + 1e-12), rtol=1e-12)


def test_conduction_residual_matches_oracle():
    import torch
    import feax_amd as fe
    from feax_amd.assembler import residual
    from oracle import hex8
    prob, bc, _ = _problem("conduction")
    m = _oracle(prob)
    ts = fe.TracedStructure.from_problem(prob)
    param = _param("conduction", prob, "node")
    rng = np.random.default_rng(4)
    T = rng.standard_normal(m.n)
    q = rng.uniform(0.2, 0.6, (len(prob.boundary_inds_list[0]), 4))
    tp = fe.TracedParams(volume_vars=(param,), surface_vars=[(q,)])
    print("conduction residual", flush=True)
    r = _np(residual(prob, ts, tp, torch.as_tensor(T, device="cuda"), bc))
    ref = hex8.apply_bc_res(m.residual(T, param, fluxes=[q]), T, bc.bc_rows, bc.bc_vals)
    err = np.abs(r - ref).max() / np.abs(ref).max()
    print("residual rel err", err)
    assert err <= 1e-13


@pytest.mark.parametrize("opts", ["cg", "direct"])
def test_conduction_solves(opts):
    """Heated bar (T = 0 and 100 on the end faces, no flux): the exact linear
    profile; with a flux on x = L: scipy's direct solve of the oracle system."""
    import scipy.sparse.linalg as spla
    import feax_amd as fe
    from oracle import hex8
    prob, _, L = _problem("conduction", dims=(12, 4, 4), flux=False)
    bc2 = fe.DirichletBCConfig([fe.DirichletBCSpec(lambda p: np.isclose(p[0], 0.0), "all", 0.0),
                                fe.DirichletBCSpec(lambda p: np.isclose(p[0], L[0]), "all", 100.0)]).create_bc(prob)
    o = fe.DirectSolverOptions(solver="cudss") if opts == "direct" else \
        fe.KrylovSolverOptions(solver="cg", tol=1e-12, atol=0.0, use_jacobi_preconditioner=True)
    k = fe.TracedParams.create_cell_var(prob, 1.7)
    print("conduction bar", opts, flush=True)
    T = np.asarray(fe.create_solver(prob, bc2, o, linear=True)(fe.TracedParams(volume_vars=(k,))))
    assert np.abs(T - 100.0 * prob.points[:, 0] / L[0]).max() <= 1e-8 * 100.0
    prob, bc, _ = _problem("conduction", dims=(12, 4, 4))
    m = _oracle(prob)
    param = _param("conduction", prob, "cell")
    tp = fe.TracedParams(volume_vars=(param,), surface_vars=[(fe.TracedParams.create_uniform_surface_var(prob, 0.8),)])
    T = np.asarray(fe.create_solver(prob, bc, o, linear=True)(tp))
    mask = np.zeros(m.n, bool)
    mask[bc.bc_rows] = True
    K = m.scipy_csr(hex8.apply_bc_csr(m.csr_values(param), m.csr, mask, bc.bc_rows, True))
    f = -m.residual(np.zeros(m.n), param, fluxes=[0.8])
    f[bc.bc_rows] = 0.0
    Td = spla.spsolve(K.tocsc(), f)
    err = np.linalg.norm(T - Td) / np.linalg.norm(Td)
    print(opts, "flux solve rel err vs scipy", err)
    assert err < 1e-9


@pytest.mark.parametrize("kind,loc", [("simp", "cell"), ("conduction", "node")])
def test_conduction_gradients(kind, loc):
    """d(T.T)/dp (cell SIMP density / node conductivity) and d(T.T)/dq per
    face point against central differences; the fused thermal compliance
    sensitivity against the autograd compliance gradient."""
    import torch
    import feax_amd as fe
    prob, bc, _ = _problem(kind, dims=(6, 3, 3))
    p0 = _param(kind, prob, loc)
    q0 = np.random.default_rng(6).uniform(0.3, 0.7, (len(prob.boundary_inds_list[0]), 4))
    opts = fe.KrylovSolverOptions(solver="cg", tol=1e-13, atol=0.0, use_jacobi_preconditioner=True)
    solver = fe.create_solver(prob, bc, opts, linear=True, return_solution=False)

    def J(pv, qv):
        with torch.no_grad():
            T = solver(fe.TracedParams(volume_vars=(torch.as_tensor(pv, device="cuda"),),
                                       surface_vars=[(torch.as_tensor(qv, device="cuda"),)]))
            return float((T * T).sum())

    r = torch.as_tensor(p0, device="cuda").requires_grad_(True)
    qq = torch.as_tensor(q0, device="cuda").requires_grad_(True)
    print("conduction gradients", kind, loc, flush=True)
    T = solver(fe.TracedParams(volume_vars=(r,), surface_vars=[(qq,)]))
    (T * T).sum().backward()
    g, gq = _np(r.grad), _np(qq.grad)
    rng = np.random.default_rng(9)
    for i in rng.choice(len(p0), 5, replace=False):
        eps = 1e-5
        pp, pm = p0.copy(), p0.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (J(pp, q0) - J(pm, q0)) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * np.abs(g).max(), (i, fd, g[i])
    for f in rng.choice(len(q0), 2, replace=False):
        for qi in (0, 3):
            qp, qm = q0.copy(), q0.copy()
            qp[f, qi] += 1e-5
            qm[f, qi] -= 1e-5
            fd = (J(p0, qp) - J(p0, qm)) / 2e-5
            assert abs(fd - gq[f, qi]) <= 1e-6 * np.abs(gq).max(), (f, qi, fd, gq[f, qi])
    if loc == "cell":
        from feax_amd.gene import compliance_and_sensitivity, create_compliance_fn
        qu = fe.TracedParams.create_uniform_surface_var(prob, 0.5)
        tp = fe.TracedParams(volume_vars=(torch.as_tensor(p0, device="cuda"),), surface_vars=[(qu,)])
        with torch.no_grad():
            Ts = solver(tp)
        C, dC = compliance_and_sensitivity(prob, tp, Ts, surface_load_params=0.5)
        r2 = torch.as_tensor(p0, device="cuda").requires_grad_(True)
        T2 = solver(fe.TracedParams(volume_vars=(r2,), surface_vars=[(qu,)]))
        create_compliance_fn(prob, surface_load_params=0.5)(T2).backward()
        ref = _np(r2.grad)
        assert np.abs(_np(dC) - ref).max() <= 1e-8 * np.abs(ref).max()


def test_thermal_solution_chains_into_the_next_solve():
    """The reference's tests/test_params_bridge.py:128-170 pattern: a solved
    temperature becomes a node variable of the next solve on the same mesh
    (here the conductivity of a second thermal solve), and the gradient of
    the second solution flows back into the first solve's conductivity."""
    import torch
    import feax_amd as fe
    prob, _, L = _problem("conduction", dims=(4, 3, 3), flux=False)
    bc = fe.DirichletBCConfig([fe.DirichletBCSpec(lambda p: np.isclose(p[0], 0.0), "all", 1.0),
                               fe.DirichletBCSpec(lambda p: np.isclose(p[0], L[0]), "all", 2.0)]).create_bc(prob)
    opts = fe.KrylovSolverOptions(solver="cg", tol=1e-13, atol=0.0, use_jacobi_preconditioner=True)
    s1 = fe.create_solver(prob, bc, opts, linear=True, return_solution=False)
    k1 = torch.full((prob.num_cells,), 1.3, dtype=torch.float64, device="cuda", requires_grad=True)
    T1 = s1(fe.TracedParams(volume_vars=(k1,)), fe.zero_like_initial_guess(prob, bc))
    Tn = fe.TracedParams.node_var_from_solution(prob, T1)
    assert tuple(Tn.shape) == (prob.num_total_nodes,) and abs(float(Tn.max()) - 2.0) < 1e-12
    bc2 = fe.DirichletBCConfig([fe.DirichletBCSpec(lambda p: np.isclose(p[0], 0.0), "all", 0.0),
                                fe.DirichletBCSpec(lambda p: np.isclose(p[0], L[0]), "all", 10.0)]).create_bc(prob)
    s2 = fe.create_solver(prob, bc2, opts, linear=True, return_solution=False)
    T2 = s2(fe.TracedParams(volume_vars=(Tn,)), fe.zero_like_initial_guess(prob, bc2))
    # k2 = T1 grows along x: the second profile bends towards the cold end
    mid = np.isclose(prob.points[:, 0], 2.0)
    assert float(T2[torch.as_tensor(np.flatnonzero(mid), device="cuda")].mean()) > 5.0
    (T2 * T2).sum().backward()
    g = _np(k1.grad)
    # T1 is homogeneous of degree 0 in k1 (a common factor leaves it
    # unchanged): sum_e k1_e dL/dk1_e = 0, while single cells matter
    assert abs(float((g * 1.3).sum())) <= 1e-8 * float(np.abs(g * 1.3).sum()) and np.abs(g).max() > 0

    def L_of(kv):
        with torch.no_grad():
            t1 = s1(fe.TracedParams(volume_vars=(torch.as_tensor(kv, device="cuda"),)),
                    fe.zero_like_initial_guess(prob, bc))
            t2 = s2(fe.TracedParams(volume_vars=(fe.TracedParams.node_var_from_solution(prob, t1),)),
                    fe.zero_like_initial_guess(prob, bc2))
            return float((t2 * t2).sum())
    k0 = np.full(prob.num_cells, 1.3)
    for e in (0, prob.num_cells // 2):
        kp, km = k0.copy(), k0.copy()
        kp[e] += 1e-5
        km[e] -= 1e-5
        fd = (L_of(kp) - L_of(km)) / 2e-5
        assert abs(fd - g[e]) <= 1e-6 * np.abs(g).max(), (e, fd, g[e])
