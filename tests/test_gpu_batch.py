"""GPU parity of batched uneven instances (segmented PCG over a
block-diagonal concatenation) against per-instance FP64 oracle solves."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _simp_instance(dims, rng):
    import feax_amd as fe
    nx = dims[0]

    class Simp(fe.Problem):
        def get_tensor_map(self):
            return fe.materials.SIMPLinearElastic(nu=0.3, E0=1.0, Emin=1e-9, penal=3.0)

        def get_surface_maps(self):
            return [fe.materials.SurfaceTraction((0.0, 0.0, 1.0))]

    prob = Simp(fe.StructuredGrid(dims, spacing=(1.0, 1.0, 1.0)).to_mesh(), vec=3, dim=3,
                location_fns=[lambda p: np.isclose(p[0], nx)])
    bc = fe.DirichletBCConfig([fe.DirichletBCSpec(lambda p: np.isclose(p[0], 0.0), "all", 0.0)]).create_bc(prob)
    rho = rng.uniform(0.3, 1.0, prob.num_cells)
    t = 10.0 ** rng.uniform(-4, -2)
    tp = fe.TracedParams(volume_vars=(rho,),
                         surface_vars=[(fe.TracedParams.create_uniform_surface_var(prob, t),)])
    return prob, bc, tp, rho, t


def _oracle_solution(prob, bc, rho, t):
    import scipy.sparse.linalg as spla
    from oracle import hex8
    m = hex8.Model(prob.points, prob.cells.astype(np.int64), prob.boundary_inds_list)
    mat = hex8.Material("simp", nu=0.3, E0=1.0, Emin=1e-9, penal=3.0)
    K = m.scipy_csr(hex8.apply_bc_csr(m.csr_values(np.zeros(m.n), rho, mat), m.csr, bc.bc_mask, bc.bc_rows))
    f = -m.residual(np.zeros(m.n), rho, mat, [hex8.constant_traction(prob.boundary_inds_list[0], t)])
    f[bc.bc_rows] = 0.0
    return spla.spsolve(K.tocsc(), f)


@pytest.fixture(scope="module")
def batch():
    from feax_amd.batch import BatchedProblem, batch_bc, batch_params
    rng = np.random.default_rng(20261203)
    # the last instance spans several vector blocks (> 4096 rows)
    dims = [(6, 3, 3), (9, 2, 4), (5, 4, 2), (12, 3, 3), (4, 2, 2), (30, 8, 8)]
    inst = [_simp_instance(d, rng) for d in dims]
    bp = BatchedProblem([i[0] for i in inst])
    bc = batch_bc(bp, [i[1] for i in inst])
    tp = batch_params(bp, [i[2] for i in inst])
    ref = [_oracle_solution(p, b, rho, t) for p, b, _, rho, t in inst]
    return dict(bp=bp, bc=bc, tp=tp, inst=inst, ref=ref)


def _errors(bp, u, ref):
    return [np.linalg.norm(np.asarray(ui.detach().cpu()) - r) / np.linalg.norm(r) for ui, r in zip(bp.split(u), ref)]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batched_solve_matches_instance_oracles(batch, precision):
    import feax_amd as fe
    from feax_amd.batch import create_batched_solver
    opts = fe.KrylovSolverOptions(solver="cg", tol=1e-10, atol=0.0, maxiter=20000, use_jacobi_preconditioner=True,
                                  precision=precision)
    solver = create_batched_solver(batch["bp"], batch["bc"], opts)
    u = solver(batch["tp"])
    errs = _errors(batch["bp"], u, batch["ref"])
    print(precision, solver.stats, "errors", ["%.1e" % e for e in errs])
    assert max(errs) < 1e-8
    assert max(solver.stats.relres) <= 2e-10
    assert solver.stats.lanes in (8, 16, 32)
    assert solver.stats.spmv["threads"] in (256, 512)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_every_spmv_configuration_matches_oracles_and_map_is_bitwise(batch, precision):
    """Every SpMV configuration the autotuner chooses from (lanes 8/16/32 x
    256/512 threads per workgroup x block-column map off/on) solves every
    instance to the oracle; for one (lanes, threads) the map run is bitwise
    the bcol run (the same products and sums in the same order, x only read
    from the LDS copy), so a wrong map entry cannot hide."""
    import feax_amd as fe
    from feax_amd.batch import create_batched_solver
    opts = fe.KrylovSolverOptions(solver="cg", tol=1e-10, atol=0.0, maxiter=20000, use_jacobi_preconditioner=True,
                                  precision=precision)
    for lanes in (8, 16, 32):
        for threads in (256, 512):
            out = {}
            for use_map in (False, True):
                cfg = {"lanes": lanes, "threads": threads, "map": use_map}
                print("case:", precision, cfg, flush=True)
                solver = create_batched_solver(batch["bp"], batch["bc"], opts, autotune=False, spmv=cfg)
                u = solver(batch["tp"])
                assert solver.stats.spmv == cfg
                errs = _errors(batch["bp"], u, batch["ref"])
                assert max(errs) < 1e-8, (cfg, errs)
                out[use_map] = (u.cpu().numpy().tobytes(), list(solver.stats.iters))
            assert out[True] == out[False], (lanes, threads)


def test_map_falls_back_to_bcol_on_wide_column_spans():
    """A scrambled node numbering (75,645 nodes in random order) spreads
    each workgroup's column nodes over more than the map's 65536-node
    bitmap: a run asked for the map reports it off and is bitwise the bcol
    run, and it solves the problem (FP64 residual from the assembly)."""
    import feax_amd as fe
    from feax_amd.batch import BatchedProblem, batch_bc, batch_params, create_batched_solver
    rng = np.random.default_rng(5)
    base, _, _, _, _ = _simp_instance((44, 40, 40), rng)
    perm = rng.permutation(base.num_total_nodes)
    pts = np.empty_like(base.points)
    pts[perm] = base.points

    class Simp(fe.Problem):
        def get_tensor_map(self):
            return fe.materials.SIMPLinearElastic(nu=0.3, E0=1.0, Emin=1e-9, penal=3.0)

        def get_surface_maps(self):
            return [fe.materials.SurfaceTraction((0.0, 0.0, 1.0))]

    prob = Simp(fe.Mesh(pts, perm[base.cells]), vec=3, dim=3, location_fns=[lambda p: np.isclose(p[0], 44.0)])
    bc = fe.DirichletBCConfig([fe.DirichletBCSpec(lambda p: np.isclose(p[0], 0.0), "all", 0.0)]).create_bc(prob)
    tp = fe.TracedParams(volume_vars=(rng.uniform(0.3, 1.0, prob.num_cells),),
                         surface_vars=[(fe.TracedParams.create_uniform_surface_var(prob, 1e-3),)])
    bp = BatchedProblem([prob])
    bcb, tpb = batch_bc(bp, [bc]), batch_params(bp, [tp])
    opts = fe.KrylovSolverOptions(solver="cg", tol=1e-8, atol=0.0, maxiter=20000, use_jacobi_preconditioner=True)
    out = {}
    for use_map in (True, False):
        print("case: scrambled numbering, map asked", use_map, flush=True)
        solver = create_batched_solver(bp, bcb, opts, autotune=False,
                                       spmv={"lanes": 16, "threads": 256, "map": use_map})
        u = solver(tpb)
        assert solver.stats.spmv["map"] is False
        assert max(solver.stats.relres) <= 1.5e-8
        out[use_map] = u.cpu().numpy().tobytes()
    assert out[True] == out[False]


def test_batched_iterations_match_single_instance_solves(batch):
    """Per-instance stopping: each instance takes (to rounding) the iterations
    its own solve takes."""
    import feax_amd as fe
    from feax_amd.batch import create_batched_solver
    opts = fe.KrylovSolverOptions(solver="cg", tol=1e-10, atol=0.0, maxiter=20000, use_jacobi_preconditioner=True)
    solver = create_batched_solver(batch["bp"], batch["bc"], opts)
    solver(batch["tp"])
    for (p, bc, tp, _, _), it in zip(batch["inst"], solver.stats.iters):
        single = fe.create_solver(p, bc, opts, linear=True, return_solution=False)
        single(tp)
        assert abs(single.stats.iters - it) <= 2, (single.stats.iters, it)


def test_batched_graph_replay_equals_eager(batch):
    """The HIP-graph replay of the iteration chunk gives bitwise the eager
    launches' result (several replays, eager work in between)."""
    import feax_amd as fe
    from feax_amd.batch import create_batched_solver
    opts = fe.KrylovSolverOptions(solver="cg", tol=1e-10, atol=0.0, maxiter=20000, use_jacobi_preconditioner=True,
                                  check_every=8)
    g = create_batched_solver(batch["bp"], batch["bc"], opts, graph=True, autotune=False)
    e = create_batched_solver(batch["bp"], batch["bc"], opts, graph=False, autotune=False)
    u1 = g(batch["tp"]).cpu().numpy()
    u2 = e(batch["tp"]).cpu().numpy()
    u3 = g(batch["tp"]).cpu().numpy()
    assert u1.tobytes() == u2.tobytes() == u3.tobytes()
    assert g.stats.iters == e.stats.iters


def test_same_mesh_parameter_batch():
    """The vmap use of feax: one mesh, several parameter sets."""
    import feax_amd as fe
    from feax_amd.batch import BatchedProblem, batch_bc, batch_params, create_batched_solver
    rng = np.random.default_rng(3)
    p, bc, tp, rho, t = _simp_instance((6, 3, 3), rng)
    rhos = [rng.uniform(0.3, 1.0, p.num_cells) for _ in range(3)]
    tps = [fe.TracedParams(volume_vars=(r,), surface_vars=tp.surface_vars) for r in rhos]
    bp = BatchedProblem([p, p, p])
    solver = create_batched_solver(bp, batch_bc(bp, [bc] * 3), fe.KrylovSolverOptions(
        solver="cg", tol=1e-10, atol=0.0, maxiter=20000, use_jacobi_preconditioner=True))
    u = solver(batch_params(bp, tps))
    for ui, r in zip(bp.split(u), rhos):
        ref = _oracle_solution(p, bc, r, t)
        assert np.linalg.norm(np.asarray(ui.cpu()) - ref) / np.linalg.norm(ref) < 1e-8


def test_recorded_directions_reproduce_each_instance_cg_solution(batch):
    """The Galerkin projection of b onto the directions recorded while CG
    solved A x = b is that CG solution (per instance; FP32 storage of the
    directions limits the agreement), and instances whose right-hand side is
    zero project to zero.  A is the plain FP32 matrix (exactly symmetric);
    the difference form's diagonal blocks carry an antisymmetric rounding
    part of order eps32 |K|, and CG run deep (rel 1e-10) on small instances
    loses the global conjugacy this identity needs."""
    import torch
    import feax_amd as fe
    from feax_amd.assembler import assemble_csr, residual
    from feax_amd.batch import SegKrylovRecord, pcg_batched
    bp, bc, tp = batch["bp"], batch["bc"], batch["tp"]
    ts = fe.TracedStructure.from_problem(bp)
    tpd = fe.TracedParams(volume_vars=(torch.as_tensor(tp.volume_vars[0], device="cuda"),),
                          surface_vars=[(torch.as_tensor(tp.surface_vars[0][0], device="cuda"),)])
    A, dinv = assemble_csr(bp, ts, tpd, None, bc, True, torch.float32, jacobi_shift=1e-12)
    b = residual(bp, ts, tpd, torch.zeros(ts.n, dtype=torch.float64, device="cuda"), bc, negate=True)
    cap = 1000
    rec = SegKrylovRecord(bp, ts.n, cap, "cuda")
    print("batched recording solve", flush=True)
    x, iters, _, _ = pcg_batched(bp, A, b, None, dinv, 1e-10, 0.0, 5000, record=rec)
    assert np.array_equal(rec.counts, np.minimum(iters, cap)) and rec.counts.min() > 0 and iters.max() <= cap
    x0 = rec.project(b)
    P, pq, bn = rec.P.cpu().numpy().astype(np.float64), rec.pq.cpu().numpy(), b.cpu().numpy()
    rows = np.concatenate([[0], np.cumsum(3 * np.diff(bp.seg_node_ptr))])
    for s, (xs, ps) in enumerate(zip(bp.split(x), bp.split(x0))):
        # the kernels against the same projection evaluated here from the record
        K, sl = int(rec.counts[s]), slice(rows[s], rows[s + 1])
        want = ((P[:K, sl] @ bn[sl]) / pq[:K, s]) @ P[:K, sl]
        got = ps.cpu().numpy()
        assert np.abs(got - want).max() <= 1e-10 * np.abs(want).max()
        # exact arithmetic: the projection is the CG solution; here FP32
        # directions and finite-precision conjugacy leave ~1e-4
        err = float((ps - xs).norm() / xs.norm())
        print("instance", s, "iters", int(iters[s]), "projection vs CG solution", err)
        assert err < 1e-3
    # zero right-hand side on one instance -> zero projection there
    row_seg = np.repeat(np.arange(bp.nseg), 3 * np.diff(bp.seg_node_ptr))
    mask = torch.as_tensor((row_seg != 2).astype(np.float64), device="cuda")
    x2 = rec.project(b * mask)
    assert float(bp.split(x2)[2].abs().max()) == 0.0
    assert float((bp.split(x2)[0] - bp.split(x0)[0]).abs().max()) == 0.0


def test_batched_mixed_recycling_cuts_inner_iterations(batch):
    """The mixed batched solve with per-instance recycling agrees with the
    one without and takes fewer inner iterations after the first outer
    step."""
    import torch
    import feax_amd as fe
    from feax_amd.batch import create_batched_solver
    bp, bc, tp = batch["bp"], batch["bc"], batch["tp"]
    tpd = fe.TracedParams(volume_vars=(torch.as_tensor(tp.volume_vars[0], device="cuda"),),
                          surface_vars=[(torch.as_tensor(tp.surface_vars[0][0], device="cuda"),)])
    out = {}
    for rec in (0, 3000):
        opts = fe.KrylovSolverOptions(solver="cg", tol=1e-10, atol=0.0, maxiter=20000, use_jacobi_preconditioner=True,
                                      precision="mixed", recycle=rec)
        s = create_batched_solver(bp, bc, opts)
        print("batched mixed, recycle", rec, flush=True)
        u = s(tpd)
        out[rec] = (u, list(s.stats.iters), s.stats.outer, max(s.stats.relres))
        print("recycle", rec, "iters", out[rec][1], "outer", out[rec][2], "relres", out[rec][3])
        assert out[rec][3] <= 1e-10
    errs = _errors(bp, out[3000][0], batch["ref"])
    assert max(errs) < 1e-8
    assert sum(out[3000][1]) < sum(out[0][1])


def test_vmap_solver_matches_one_solve_per_lane():
    """create_vmap_solver over a leading batch axis of the traced variables
    and of the Dirichlet values (the vmap uses of the reference's
    tests/test_vmap.py and tests/test_bc_vmap.py) equals one create_solver
    call per lane."""
    import torch
    import feax_amd as fe
    from feax_amd.batch import create_vmap_solver
    rng = np.random.default_rng(11)
    p, bc, tp, rho, t = _simp_instance((6, 3, 3), rng)
    B = 3
    rhos = rng.uniform(0.3, 1.0, (B, p.num_cells))
    ts_ = np.stack([fe.TracedParams.create_uniform_surface_var(p, 10.0 ** rng.uniform(-3, -2)) for _ in range(B)])
    vals = np.stack([bc.vals_numpy() + 0.01 * (i + 1) * (np.asarray(bc.bc_rows) % 3 == 2) for i in range(B)])
    opts = fe.KrylovSolverOptions(solver="cg", tol=1e-10, atol=0.0, maxiter=20000, use_jacobi_preconditioner=True)
    solve = create_vmap_solver(p, bc, opts)
    print("vmap solve", flush=True)
    U = solve(fe.TracedParams(volume_vars=(torch.as_tensor(rhos, device="cuda"),),
                              surface_vars=[(torch.as_tensor(ts_, device="cuda"),)]), bc_vals=vals)
    assert tuple(U.shape) == (B, p.num_total_dofs_all_vars)
    single = fe.create_solver(p, bc, opts, linear=True, return_solution=False)
    for i in range(B):
        bci = bc.replace_vals(vals[i])
        ui = single(fe.TracedParams(volume_vars=(rhos[i],), surface_vars=[(ts_[i],)]),
                    fe.zero_like_initial_guess(p, bci), bc=bci)
        err = float((U[i] - ui).norm() / ui.norm())
        print("lane", i, "vs single solve", err)
        assert err < 1e-8


def test_vmap_solver_gradients_match_per_lane_autograd():
    """grad of sum_i w_i |u_i|^2 through the vmapped solve (the reference's
    vmap-of-grad composition, tests/test_vmap.py:243-290) equals each lane's
    gradient through its own create_solver call, for the cell densities
    (B, E) and the per-face-point tractions (B, F, 4)."""
    import torch
    import feax_amd as fe
    from feax_amd.batch import create_vmap_solver
    rng = np.random.default_rng(12)
    p, bc, tp, rho, t = _simp_instance((6, 3, 3), rng)
    B = 3
    rhos = rng.uniform(0.3, 1.0, (B, p.num_cells))
    ts_ = np.stack([fe.TracedParams.create_uniform_surface_var(p, 10.0 ** rng.uniform(-3, -2)) for _ in range(B)])
    ts_ = ts_ * rng.uniform(0.5, 1.5, ts_.shape)
    w = torch.tensor([1.0, 2.0, 0.5], dtype=torch.float64, device="cuda")
    opts = fe.KrylovSolverOptions(solver="cg", tol=1e-12, atol=0.0, maxiter=20000, use_jacobi_preconditioner=True)
    r = torch.tensor(rhos, device="cuda", requires_grad=True)
    s = torch.tensor(ts_, device="cuda", requires_grad=True)
    print("vmap gradient", flush=True)
    U = create_vmap_solver(p, bc, opts)(fe.TracedParams(volume_vars=(r,), surface_vars=[(s,)]))
    (w * (U * U).sum(dim=1)).sum().backward()
    single = fe.create_solver(p, bc, opts, linear=True, return_solution=False)
    for i in range(B):
        ri = torch.tensor(rhos[i], device="cuda", requires_grad=True)
        si = torch.tensor(ts_[i], device="cuda", requires_grad=True)
        ui = single(fe.TracedParams(volume_vars=(ri,), surface_vars=[(si,)]))
        (w[i] * (ui * ui).sum()).backward()
        er = float((r.grad[i] - ri.grad).norm() / ri.grad.norm())
        es = float((s.grad[i] - si.grad).norm() / si.grad.norm())
        print("lane", i, "rho grad", er, "traction grad", es)
        assert er < 1e-8 and es < 1e-8


@pytest.mark.parametrize("method", ["bicgstab", "gmres"])
def test_vmap_solver_other_krylov_methods(method):
    """The reference's tests/test_vmap.py:109-190 cases: vmapped BiCGSTAB and
    GMRES solves equal one solve per lane, and their gradients too."""
    import torch
    import feax_amd as fe
    from feax_amd.batch import create_vmap_solver
    rng = np.random.default_rng(13)
    p, bc, tp, rho, t = _simp_instance((6, 3, 3), rng)
    B = 2
    rhos = rng.uniform(0.3, 1.0, (B, p.num_cells))
    ts_ = np.stack([fe.TracedParams.create_uniform_surface_var(p, 1e-2) for _ in range(B)])
    opts = fe.KrylovSolverOptions(solver=method, tol=1e-12, atol=0.0, maxiter=20000, use_jacobi_preconditioner=True)
    r = torch.tensor(rhos, device="cuda", requires_grad=True)
    print("vmap", method, flush=True)
    U = create_vmap_solver(p, bc, opts)(fe.TracedParams(volume_vars=(r,), surface_vars=[(torch.as_tensor(ts_),)]))
    (U * U).sum().backward()
    single = fe.create_solver(p, bc, opts, linear=True, return_solution=False)
    for i in range(B):
        ri = torch.tensor(rhos[i], device="cuda", requires_grad=True)
        ui = single(fe.TracedParams(volume_vars=(ri,), surface_vars=[(ts_[i],)]))
        (ui * ui).sum().backward()
        assert float((U[i] - ui).norm() / ui.norm()) < 1e-12
        assert float((r.grad[i] - ri.grad).norm() / ri.grad.norm()) < 1e-12
