"""Upper-block storage of the FP32 difference form (fx_b3sym_compact_f32,
fx_csr_spmv_b3sym_f32f64, fx_pcg_b3sym_f32f64): the mixed form's PCG runs its
SpMV on the upper 3x3 blocks only, reading each lower block as the transpose
of its mirror.  Checked against the full difference-form kernel (same
operator, different summation grouping: FP64 rounding only), against the
block-column-map PCG, and for the refusal of a non-symmetric matrix."""

import ctypes

import numpy as np
import pytest

from conftest import cantilever

pytestmark = pytest.mark.gpu

EPS64 = np.finfo(np.float64).eps


def _np(t):
    return t.detach().cpu().numpy()


def _distorted_problem(dims=(9, 5, 4), seed=8):
    """A distorted, renumbered box (general upper / lower block structure)."""
    import torch
    import feax_amd as fe
    rng = np.random.default_rng(seed)
    mesh = fe.StructuredGrid(dims, spacing=(1.0, 1.0, 1.0)).to_mesh()
    pts = mesh.points + rng.uniform(-0.1, 0.1, mesh.points.shape)
    perm = rng.permutation(len(pts))
    cells = np.argsort(perm)[mesh.cells][rng.permutation(len(mesh.cells))]
    pts = pts[perm]
    L = float(dims[0])

    class P(fe.Problem):
        def get_tensor_map(self):
            return fe.materials.IsotropicLinearElastic(nu=0.3)

        def get_surface_maps(self):
            return [fe.materials.SurfaceTraction((0.0, 0.0, 1.0))]

    prob = P(fe.Mesh(pts, cells), vec=3, dim=3, location_fns=[lambda p: np.isclose(p[0], L, atol=0.2)])
    bc = fe.DirichletBCConfig([fe.DirichletBCSpec(lambda p: p[0] < 0.2, "all", 0.0)]).create_bc(prob)
    ts = fe.TracedStructure.from_problem(prob)
    E = torch.as_tensor(rng.uniform(1.0, 2.0, prob.num_cells), device="cuda")
    tp = fe.TracedParams(volume_vars=(E,), surface_vars=[(fe.TracedParams.create_uniform_surface_var(prob, 0.1),)])
    return prob, bc, ts, tp


def test_sym_spmv_equals_the_difference_form_kernel():
    import torch
    from feax_amd import _native as nat
    from feax_amd.assembler import assemble_csr
    prob, bc, ts, tp = _distorted_problem()
    print("case: b3sym spmv, distorted 9x5x4", flush=True)
    A, _ = assemble_csr(prob, ts, tp, None, bc, True, torch.float32, diff_form=True, jacobi_shift=1e-12)
    A64 = assemble_csr(prob, ts, tp, None, bc, True, torch.float64)
    up = A.upper_storage()
    assert up is not None
    sym, vals = up
    nbl = A.nse // 9
    assert sym.num_upper * 2 > nbl and sym.num_upper < nbl          # about half the blocks
    rng = np.random.default_rng(3)
    x = torch.as_tensor(1.0 + 1e-2 * rng.standard_normal(ts.n), device="cuda")
    y_full = A.matvec(x)
    y_sym = torch.empty_like(y_full)
    nat.call("fx_csr_spmv_b3sym_f32f64", nat.stream(), ts.n, nat.ptr(A.indptr), ctypes.addressof(sym.struct),
             nat.ptr(vals), nat.ptr(x), nat.ptr(y_sym))
    import scipy.sparse as sp
    K = sp.csr_matrix((_np(A64.data), _np(A64.indices), _np(A64.indptr)), shape=(ts.n, ts.n))
    scale = abs(K) @ np.abs(_np(x))
    err = np.abs(_np(y_sym) - _np(y_full))
    print("sym vs diff3 max err / (|K||x| eps)", (err / (scale * EPS64)).max())
    assert np.all(err <= 64 * EPS64 * scale)
    # the compact copy is exactly the upper blocks of the full storage
    tab_up = _np(sym.up_ptr)
    blocks = _np(A.data).reshape(-1, 9)
    node = np.repeat(np.arange(ts.n // 3), np.diff(_np(A.indptr)[0::3]) // 9)   # 9 deg values per node
    upper_mask = _np(A.bcol) >= node
    assert np.array_equal(_np(vals).reshape(-1, 9), blocks[upper_mask])
    assert tab_up[-1] == sym.num_upper


def test_sym_spmv_with_more_than_512_columns_per_workgroup():
    """A 4 x 4 x 100 box in natural numbering (the last axis fastest): a
    64-node workgroup lies on one line of 101 nodes and reaches 594 unique
    column nodes, more than its 512 threads, so the SpMV stages the rest in
    its second loop.  The upper-block product equals the full-storage
    difference form."""
    import torch
    import feax_amd as fe
    from feax_amd import _native as nat
    from feax_amd.assembler import assemble_csr
    print("case: b3sym spmv, 4x4x100 box, > 512 columns per workgroup", flush=True)
    mesh = fe.StructuredGrid((4, 4, 100), spacing=(1.0, 1.0, 1.0)).to_mesh()

    class P(fe.Problem):
        def get_tensor_map(self):
            return fe.materials.IsotropicLinearElastic(nu=0.3)

    prob = P(mesh, vec=3, dim=3)
    bc = fe.DirichletBCConfig([fe.DirichletBCSpec(lambda p: p[0] < 0.5, "all", 0.0)]).create_bc(prob)
    ts = fe.TracedStructure.from_problem(prob)
    E = torch.ones(prob.num_cells, dtype=torch.float64, device="cuda")
    tp = fe.TracedParams(volume_vars=(E,))
    A, _ = assemble_csr(prob, ts, tp, None, bc, True, torch.float32, diff_form=True, jacobi_shift=1e-12)
    A64 = assemble_csr(prob, ts, tp, None, bc, True, torch.float64)
    up = A.upper_storage()
    assert up is not None
    sym, vals = up
    assert sym.struct.max_unique > 512
    rng = np.random.default_rng(5)
    x = torch.as_tensor(1.0 + 1e-2 * rng.standard_normal(ts.n), device="cuda")
    y_full = A.matvec(x)
    y_sym = torch.empty_like(y_full)
    nat.call("fx_csr_spmv_b3sym_f32f64", nat.stream(), ts.n, nat.ptr(A.indptr), ctypes.addressof(sym.struct),
             nat.ptr(vals), nat.ptr(x), nat.ptr(y_sym))
    import scipy.sparse as sp
    K = sp.csr_matrix((_np(A64.data), _np(A64.indices), _np(A64.indptr)), shape=(ts.n, ts.n))
    scale = abs(K) @ np.abs(_np(x))
    err = np.abs(_np(y_sym) - _np(y_full))
    print("max err / (|K||x| eps)", (err / (scale * EPS64)).max(), "max_unique", sym.struct.max_unique)
    assert np.all(err <= 64 * EPS64 * scale)


def test_fp32_arithmetic_assembly_matches_the_fp64_formed_one():
    """FX_ASM_COMPUTE32: element matrices formed in FP32 (both passes of the
    two-pass assembly in their float instantiations) agree with the FP64-formed
    FP32 matrix to FP32 rounding of the node's largest value, plain CSR and
    difference form, and are bitwise repeatable."""
    import torch
    from feax_amd.assembler import assemble_csr
    prob, bc, ts, tp = _distorted_problem()
    print("case: hex8 assembly with FP32 arithmetic", flush=True)
    for diff in (False, True):
        A, d = assemble_csr(prob, ts, tp, None, bc, True, torch.float32, diff_form=diff, jacobi_shift=1e-12)
        B, e = assemble_csr(prob, ts, tp, None, bc, True, torch.float32, diff_form=diff, jacobi_shift=1e-12,
                            compute32=True)
        B2, _ = assemble_csr(prob, ts, tp, None, bc, True, torch.float32, diff_form=diff, jacobi_shift=1e-12,
                             compute32=True)
        assert torch.equal(B.data.view(torch.int32), B2.data.view(torch.int32))
        a, b_ = _np(A.data).astype(np.float64), _np(B.data).astype(np.float64)
        # per node: its three rows' values are contiguous in both layouts
        seg = _np(A.indptr)[0::3]
        nodemax = np.repeat(np.maximum.reduceat(np.abs(a), seg[:-1]), np.diff(seg))
        rel = np.abs(a - b_) / nodemax
        print("diff form" if diff else "csr", "max |FP32-formed - FP64-formed| / node max", rel.max())
        assert rel.max() < 2e-5
        assert np.allclose(_np(e), _np(d), rtol=1e-5, atol=0)
    with pytest.raises(ValueError):
        assemble_csr(prob, ts, tp, None, bc, True, torch.float64, compute32=True)


def test_sym_storage_refused_for_row_only_elimination():
    import torch
    from feax_amd import _native as nat
    from feax_amd.assembler import assemble_csr
    prob, bc, ts, tp = _distorted_problem()
    print("case: b3sym refusal, row-only elimination", flush=True)
    A, _ = assemble_csr(prob, ts, tp, None, bc, False, torch.float32, diff_form=True, jacobi_shift=1e-12)
    sym = ts.b3sym()
    assert sym is not None
    upper = torch.empty(9 * sym.num_upper, dtype=torch.float32, device="cuda")
    bad = torch.empty(2, dtype=torch.int32, device="cuda")
    nat.call("fx_b3sym_compact_f32", nat.stream(), ts.n, nat.ptr(A.indptr), nat.ptr(A.bcol),
             ctypes.addressof(sym.struct), nat.ptr(A.data), nat.ptr(upper), nat.ptr(bad))
    b = _np(bad)
    print("asymmetric blocks", b)
    assert b[0] > 0 and b[1] == 0
    assert A.upper_storage() is None


def test_sym_pcg_matches_the_map_pcg_and_repeats_bitwise():
    import torch
    from feax_amd.assembler import assemble_csr, residual
    from feax_amd.solvers.common import pcg
    prob, bc, ts, tp = _distorted_problem()
    print("case: b3sym pcg vs map pcg", flush=True)
    A, dinv = assemble_csr(prob, ts, tp, None, bc, True, torch.float32, diff_form=True, jacobi_shift=1e-12)
    b = residual(prob, ts, tp, torch.zeros(ts.n, dtype=torch.float64, device="cuda"), bc, negate=True)
    x1, i1 = pcg(A, b, None, 1e-10, 0.0, 5000, dinv)                 # upper-block storage
    x1b, i1b = pcg(A, b, None, 1e-10, 0.0, 5000, dinv)
    x2, i2 = pcg(A, b, None, 1e-10, 0.0, 5000, dinv, use_map=True)   # full storage, block-column map
    print("iterations sym", i1.iters, "map", i2.iters)
    assert A.upper_storage() is not None
    assert abs(i1.iters - i2.iters) <= max(2, 0.02 * i2.iters)
    assert i1.iters == i1b.iters and _np(x1).tobytes() == _np(x1b).tobytes()
    d = np.abs(_np(x1) - _np(x2)).max() / np.abs(_np(x2)).max()
    print("rel diff", d)
    assert d < 1e-8


def test_mixed_solver_runs_the_sym_kernel_and_matches_direct():
    import scipy.sparse.linalg as spl
    import feax_amd as fe
    from oracle import hex8, krylov
    from conftest import E_MOD, NU, TRACTION, oracle_model
    from feax_amd.solvers.common import SpmvTimer
    problem, bc, tp = cantilever((50, 5, 5), (2.0, 2.0, 2.0), 100.0)
    print("case: mixed solver on b3sym, cantilever 50x5x5", flush=True)
    opts = fe.KrylovSolverOptions(solver="cg", tol=1e-8, atol=0.0, maxiter=20000, use_jacobi_preconditioner=True,
                                  precision="mixed")
    solver = fe.create_solver(problem, bc, opts, linear=True, traced_params=tp)
    with SpmvTimer() as t:
        u = np.asarray(solver(tp, fe.zero_like_initial_guess(problem, bc)))
    assert "f32f64/sym3" in t.kernels and t.kernels["f32f64/sym3"][1] > 0
    # the oracle's FP64 system, solved directly
    m = oracle_model(problem)
    mat = hex8.Material("linear_elastic", nu=NU)
    tr = [hex8.constant_traction(problem.boundary_inds_list[0], TRACTION)]
    _, info = krylov.linear_solve(m, mat, np.full(m.N, E_MOD), tr, bc.bc_rows, bc.bc_vals, tol=1e-6, atol=0.0)
    ud = spl.spsolve(info["K"].tocsc(), info["b"])
    err = np.linalg.norm(u - ud) / np.linalg.norm(ud)
    print("mixed (sym) vs direct", err, solver.stats)
    assert solver.stats.relres <= 1e-8 and err < 1e-6


@pytest.mark.parametrize("loc", ["cell", "node"])
def test_upper_assembly_equals_the_compacted_full_difference_form(loc):
    """FX_ASM_UPPER writes bitwise the array that the compaction of the full
    difference form makes; the Jacobi diagonal is unchanged."""
    import torch
    import feax_amd as fe
    from feax_amd.assembler import assemble_csr
    prob, bc, ts, tp = _distorted_problem()
    if loc == "node":
        rng = np.random.default_rng(4)
        tp = fe.TracedParams(volume_vars=(torch.as_tensor(rng.uniform(1.0, 2.0, prob.num_total_nodes), device="cuda"),),
                             surface_vars=tp.surface_vars)
    print("case: upper assembly,", loc, flush=True)
    Au, du = assemble_csr(prob, ts, tp, None, bc, True, torch.float32, diff_form=True, jacobi_shift=1e-12, upper=True)
    Af, df = assemble_csr(prob, ts, tp, None, bc, True, torch.float32, diff_form=True, jacobi_shift=1e-12)
    assert Au.form == "sym3" and Af.form == "diff3"
    sym, vals = Af.upper_storage()
    assert Au.data.numel() == vals.numel() == 9 * sym.num_upper
    assert torch.equal(Au.data.view(torch.int32), vals.view(torch.int32))
    assert torch.equal(du, df)
    x = torch.as_tensor(np.random.default_rng(2).standard_normal(ts.n), device="cuda")
    assert torch.equal(Au.matvec(x), Au.matvec(x))
    # row-only elimination cannot use it: the full difference form comes back
    Ar, _ = assemble_csr(prob, ts, tp, None, bc, False, torch.float32, diff_form=True, jacobi_shift=1e-12, upper=True)
    assert Ar.form == "diff3"


def test_gather_node_order_does_not_change_the_matrix():
    """The per-node gather's processing order (Morton, XCD-contiguous) is a
    schedule only: the assembled values are bitwise those of the natural
    order, FP64 and FP32 difference form."""
    import torch
    from feax_amd.assembler import assemble_csr, morton_order
    prob, bc, ts, tp = _distorted_problem()
    print("case: gather node order", flush=True)
    assert ts.node_order is None
    A1, d1 = assemble_csr(prob, ts, tp, None, bc, True, torch.float64, jacobi_shift=1e-12)
    B1, _ = assemble_csr(prob, ts, tp, None, bc, True, torch.float32, diff_form=True, jacobi_shift=1e-12)
    order = morton_order(prob.points)
    assert sorted(order.tolist()) == list(range(ts.num_nodes))
    ts.node_order = torch.as_tensor(order, device="cuda")
    try:
        A2, d2 = assemble_csr(prob, ts, tp, None, bc, True, torch.float64, jacobi_shift=1e-12)
        B2, _ = assemble_csr(prob, ts, tp, None, bc, True, torch.float32, diff_form=True, jacobi_shift=1e-12)
    finally:
        ts.node_order = None
    assert torch.equal(A1.data, A2.data) and torch.equal(d1, d2)
    assert torch.equal(B1.data.view(torch.int32), B2.data.view(torch.int32))


def test_fp64_sym_spmv_and_pcg_match_the_block3_path():
    """FP64 block-3 matrix on upper-block storage (32-node workgroups): the
    SpMV equals the block-3 kernel to FP64 rounding, the PCG matches the
    full-storage PCG, and a row-only eliminated matrix is refused."""
    import torch
    import scipy.sparse as sp
    from feax_amd import _native as nat
    from feax_amd.assembler import assemble_csr, residual
    from feax_amd.solvers.common import pcg, SpmvTimer
    prob, bc, ts, tp = _distorted_problem()
    print("case: fp64 b3sym", flush=True)
    A, dinv = assemble_csr(prob, ts, tp, None, bc, True, torch.float64, jacobi_shift=1e-12)
    assert A.sym_hint
    up = A.upper_storage()
    assert up is not None
    sym, vals = up
    assert sym.struct.nodes_per_block == 32 and vals.dtype == torch.float64
    x = torch.as_tensor(np.random.default_rng(6).standard_normal(ts.n), device="cuda")
    y_b3 = A.matvec(x)
    y_sym = torch.empty_like(y_b3)
    nat.call("fx_csr_spmv_b3sym_f64", nat.stream(), ts.n, nat.ptr(A.indptr), ctypes.addressof(sym.struct),
             nat.ptr(vals), nat.ptr(x), nat.ptr(y_sym))
    K = sp.csr_matrix((_np(A.data), _np(A.indices), _np(A.indptr)), shape=(ts.n, ts.n))
    scale = abs(K) @ np.abs(_np(x))
    err = np.abs(_np(y_sym) - _np(y_b3))
    print("fp64 sym vs block3 max err / (|K||x| eps)", (err / (scale * EPS64 + 1e-300)).max())
    assert np.all(err <= 64 * EPS64 * scale)
    b = residual(prob, ts, tp, torch.zeros(ts.n, dtype=torch.float64, device="cuda"), bc, negate=True)
    with SpmvTimer() as t:
        x1, i1 = pcg(A, b, None, 1e-10, 0.0, 5000, dinv, True)
    assert "f64/sym3" in t.kernels
    x2, i2 = pcg(A, b, None, 1e-10, 0.0, 5000, dinv, True, use_map=False)
    print("iterations sym", i1.iters, "block3", i2.iters)
    assert abs(i1.iters - i2.iters) <= max(2, 0.02 * i2.iters)
    assert np.abs(_np(x1) - _np(x2)).max() <= 1e-8 * np.abs(_np(x2)).max()
    # row-only elimination: not symmetric, refused by the check
    Ar = assemble_csr(prob, ts, tp, None, bc, False, torch.float64)
    assert not Ar.sym_hint and Ar.upper_storage() is None
    upper = torch.empty(9 * sym.num_upper, dtype=torch.float64, device="cuda")
    bad = torch.empty(2, dtype=torch.int32, device="cuda")
    nat.call("fx_b3sym_compact_f64", nat.stream(), ts.n, nat.ptr(Ar.indptr), nat.ptr(Ar.bcol),
             ctypes.addressof(sym.struct), nat.ptr(Ar.data), nat.ptr(upper), nat.ptr(bad))
    bb = _np(bad)
    assert bb[0] > 0 and bb[1] == 0
