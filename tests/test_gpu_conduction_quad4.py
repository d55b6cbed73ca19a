"""GPU parity of the scalar QUAD4 kernels (plane conduction, vec = 1)
against oracle/conduction.py:QuadModel: assembly, the residual with a
per-point edge flux, CG and direct solves against scipy, and gradients of
the conductivity (node variable) and the flux against central differences."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy()


def _problem(n=(10, 5), L=(2.0, 1.0), seed=0):
    import feax_amd as fe
    from oracle import quad4
    pts, cells = quad4.rectangle(n[0], n[1], L[0], L[1])
    rng = np.random.default_rng(seed)
    inner = (pts[:, 0] > 1e-9) & (pts[:, 0] < L[0] - 1e-9) & (pts[:, 1] > 1e-9) & (pts[:, 1] < L[1] - 1e-9)
    pts = pts.copy()
    pts[inner] += rng.uniform(-0.15, 0.15, (inner.sum(), 2)) * min(L[0] / n[0], L[1] / n[1])
    perm = rng.permutation(len(pts))
    pts, cells = pts[perm], np.argsort(perm)[cells][rng.permutation(len(cells))]

    class Plane(fe.Problem):
        def get_tensor_map(self):
            return fe.materials.IsotropicConduction()

        def get_surface_maps(self):
            return [fe.materials.SurfaceFlux()]

    prob = Plane(fe.Mesh(pts, cells, ele_type="QUAD4"), vec=1, dim=2, ele_type="QUAD4",
                 location_fns=[lambda p: np.isclose(p[0], L[0])])
    bc = fe.DirichletBCConfig([fe.DirichletBCSpec(lambda p: np.isclose(p[0], 0.0), "all", 0.0)]).create_bc(prob)
    return prob, bc


def test_plane_conduction_assembly_residual_and_solves():
    import scipy.sparse.linalg as spla
    import torch
    import feax_amd as fe
    from feax_amd.assembler import assemble_csr, residual
    from oracle import conduction, hex8
    prob, bc = _problem()
    m = conduction.QuadModel(prob.points, prob.cells.astype(np.int64), prob.boundary_inds_list)
    ts = fe.TracedStructure.from_problem(prob)
    rng = np.random.default_rng(1)
    k = rng.uniform(0.5, 2.0, prob.num_total_nodes)
    print("plane conduction assembly", flush=True)
    A = assemble_csr(prob, ts, fe.TracedParams(volume_vars=(k,)), None, bc, True, torch.float64)
    mask = np.zeros(m.n, bool)
    mask[bc.bc_rows] = True
    ref = hex8.apply_bc_csr(m.csr_values(k), m.csr, mask, bc.bc_rows, True)
    assert np.array_equal(_np(A.indices), m.csr["indices"])
    assert np.abs(_np(A.data) - ref).max() <= 1e-13 * np.abs(ref).max()
    T = rng.standard_normal(m.n)
    q = rng.uniform(0.2, 0.6, (len(prob.boundary_inds_list[0]), 2))
    tp = fe.TracedParams(volume_vars=(k,), surface_vars=[(q,)])
    r = _np(residual(prob, ts, tp, torch.as_tensor(T, device="cuda"), bc))
    rref = hex8.apply_bc_res(m.residual(T, k, fluxes=[q]), T, bc.bc_rows, bc.bc_vals)
    assert np.abs(r - rref).max() <= 1e-13 * np.abs(rref).max()
    K = m.scipy_csr(ref)
    f = -m.residual(np.zeros(m.n), k, fluxes=[q])
    f[bc.bc_rows] = 0.0
    Td = spla.spsolve(K.tocsc(), f)
    for o in (fe.KrylovSolverOptions(solver="cg", tol=1e-12, atol=0.0, use_jacobi_preconditioner=True),
              fe.DirectSolverOptions(solver="cudss")):
        Ts = np.asarray(fe.create_solver(prob, bc, o, linear=True)(tp))
        assert np.linalg.norm(Ts - Td) / np.linalg.norm(Td) < 1e-9


def test_plane_conduction_gradients():
    import torch
    import feax_amd as fe
    prob, bc = _problem((6, 3))
    rng = np.random.default_rng(2)
    k0 = rng.uniform(0.5, 2.0, prob.num_total_nodes)
    q0 = rng.uniform(0.2, 0.6, (len(prob.boundary_inds_list[0]), 2))
    solver = fe.create_solver(prob, bc, fe.KrylovSolverOptions(solver="cg", tol=1e-13, atol=0.0,
                              use_jacobi_preconditioner=True), linear=True, return_solution=False)

    def J(kv, qv):
        with torch.no_grad():
            T = solver(fe.TracedParams(volume_vars=(torch.as_tensor(kv, device="cuda"),),
                                       surface_vars=[(torch.as_tensor(qv, device="cuda"),)]))
            return float((T * T).sum())
    kt = torch.tensor(k0, device="cuda", requires_grad=True)
    qt = torch.tensor(q0, device="cuda", requires_grad=True)
    print("plane conduction gradients", flush=True)
    T = solver(fe.TracedParams(volume_vars=(kt,), surface_vars=[(qt,)]))
    (T * T).sum().backward()
    gk, gq = _np(kt.grad), _np(qt.grad)
    for i in rng.choice(len(k0), 4, replace=False):
        p, m = k0.copy(), k0.copy()
        p[i] += 1e-5
        m[i] -= 1e-5
        fd = (J(p, q0) - J(m, q0)) / 2e-5
        assert abs(fd - gk[i]) <= 1e-6 * np.abs(gk).max(), (i, fd, gk[i])
    for f in range(min(2, len(q0))):
        p, m = q0.copy(), q0.copy()
        p[f, 1] += 1e-5
        m[f, 1] -= 1e-5
        fd = (J(k0, p) - J(k0, m)) / 2e-5
        assert abs(fd - gq[f, 1]) <= 1e-6 * np.abs(gq).max(), (f, fd, gq[f, 1])
