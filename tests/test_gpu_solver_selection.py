"""create_solver's "auto" resolution and the failure paths of the solvers.

* detect_matrix_property on assembled J_bc (feax/solvers/options.py:188-259):
  symmetric elimination -> SPD, row-only elimination -> GENERAL, a negated
  matrix -> SYMMETRIC, an UPPER view -> symmetric by construction.
* create_solver resolves "auto" from a sample J_bc (feax/solver.py:303-373):
  cg for the linear elastic solve, gmres for a Newton solve with
  symmetric_elimination=False, whose solution equals the symmetric one.
* check_convergence NaN-fills a solution that misses the threshold
  (feax/solvers/common.py:376-395), for the FP64 and the mixed forms.
* Newton's Armijo line search: with no backtracks allowed the search fails
  by construction (a host-side setting, every kernel runs normally); feax
  then raises NewtonLineSearchError, or takes the 1e-8 fallback step when
  raise_on_line_search_failure=False (feax/solvers/newton.py:809-859,
  1004-1025).
"""

import numpy as np
import pytest

from conftest import cantilever

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy()


def _small():
    return cantilever((12, 3, 3), (1.0, 1.0, 1.0), 12.0)


def test_detect_matrix_property_on_assembled_matrices():
    import torch
    import feax_amd as fe
    from feax_amd.assembler import assemble_csr
    p, bc, tp = _small()
    ts = fe.TracedStructure.from_problem(p)
    print("case: detect_matrix_property on a 12x3x3 cantilever", flush=True)
    J = assemble_csr(p, ts, tp, None, bc, True, torch.float64)
    assert fe.detect_matrix_property(J) == fe.MatrixProperty.SPD
    Jr = assemble_csr(p, ts, tp, None, bc, False, torch.float64)
    assert fe.detect_matrix_property(Jr) == fe.MatrixProperty.GENERAL
    assert fe.detect_matrix_property(J.with_data(-J.data)) == fe.MatrixProperty.SYMMETRIC
    # a non-symmetric matrix stored as a view is symmetric by construction
    assert fe.detect_matrix_property(Jr, matrix_view=fe.MatrixView.UPPER) == fe.MatrixProperty.SPD
    # the explicit transpose of the row-only matrix is its scipy transpose
    T = Jr.T.to_scipy()
    assert abs(T - Jr.to_scipy().T).max() == 0.0


def test_auto_resolves_cg_for_the_linear_solve(capsys):
    import feax_amd as fe
    p, bc, tp = _small()
    solver = fe.create_solver(p, bc, fe.KrylovSolverOptions(tol=1e-10, atol=0.0, use_jacobi_preconditioner=True),
                              linear=True, traced_params=tp)
    out = capsys.readouterr().out
    assert "matrix_property=SPD -> cg" in out
    ref = fe.create_solver(p, bc, fe.KrylovSolverOptions(solver="cg", tol=1e-10, atol=0.0,
                                                         use_jacobi_preconditioner=True), linear=True)
    assert np.array_equal(np.asarray(solver(tp)), np.asarray(ref(tp)))


def test_auto_resolves_gmres_for_row_only_newton(capsys):
    import feax_amd as fe
    p, bc, tp = _small()
    opts = fe.KrylovSolverOptions(tol=1e-12, atol=0.0, use_jacobi_preconditioner=True, restart=50)
    print("case: auto Newton, symmetric_elimination=False", flush=True)
    # feax's default rel_tol 1e-8: one GMRES step (Jacobi, left-preconditioned)
    # reaches it.  A tighter target leaves a residual whose preconditioned
    # norm is below machine epsilon, which jax's gmres treats as zero (no
    # step), and the line search then fails as it would in feax.
    nopts = fe.NewtonOptions(rel_tol=1e-8)
    solver = fe.create_solver(p, bc, opts, newton_options=nopts, linear=False, traced_params=tp,
                              symmetric_elimination=False)
    assert "matrix_property=GENERAL -> gmres" in capsys.readouterr().out
    u = np.asarray(solver(tp))
    sym = fe.create_solver(p, bc, fe.KrylovSolverOptions(solver="cg", tol=1e-12, atol=0.0,
                                                         use_jacobi_preconditioner=True),
                           newton_options=nopts, linear=False)
    us = np.asarray(sym(tp))
    err = np.linalg.norm(u - us) / np.linalg.norm(us)
    print("row-only gmres Newton vs symmetric cg Newton", err, solver.stats, sym.stats)
    assert err < 1e-6
    # the linear path of feax ignores symmetric_elimination (always symmetric)
    lin = fe.create_solver(p, bc, fe.KrylovSolverOptions(solver="cg", tol=1e-12, atol=0.0,
                                                         use_jacobi_preconditioner=True),
                           linear=True, symmetric_elimination=False)
    assert np.linalg.norm(np.asarray(lin(tp)) - us) <= 1e-6 * np.linalg.norm(us)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_check_convergence_nan_fills_an_unconverged_solve(precision):
    import feax_amd as fe
    p, bc, tp = _small()
    print("case: check_convergence NaN fill", precision, flush=True)
    bad = fe.create_solver(p, bc, fe.KrylovSolverOptions(solver="cg", tol=1e-10, atol=0.0, maxiter=1,
                                                         use_jacobi_preconditioner=True, precision=precision,
                                                         check_convergence=True), linear=True)
    u = np.asarray(bad(tp))
    assert u.shape == (p.num_total_dofs_all_vars,) and np.all(np.isnan(u))
    good = fe.create_solver(p, bc, fe.KrylovSolverOptions(solver="cg", tol=1e-10, atol=0.0, maxiter=5000,
                                                          use_jacobi_preconditioner=True, precision=precision,
                                                          check_convergence=True), linear=True)
    assert np.all(np.isfinite(np.asarray(good(tp))))


def test_newton_line_search_failure_raises_or_takes_the_fallback_step():
    import feax_amd as fe
    from feax_amd.solvers.newton import NewtonLineSearchError
    p, bc, tp = _small()
    lin = fe.KrylovSolverOptions(solver="cg", tol=1e-12, atol=0.0, use_jacobi_preconditioner=True)
    print("case: Newton with no line-search backtracks", flush=True)
    s = fe.create_solver(p, bc, lin, newton_options=fe.NewtonOptions(line_search_max_backtracks=0), linear=False)
    with pytest.raises(NewtonLineSearchError):
        s(tp)
    s2 = fe.create_solver(p, bc, lin, newton_options=fe.NewtonOptions(line_search_max_backtracks=0,
                                                                      raise_on_line_search_failure=False),
                          linear=False)
    u = np.asarray(s2(tp))
    assert s2.stats.newton_iters == 1 and s2.stats.alphas == [1e-8]
    # linear material: the Newton direction from u0 is the exact increment
    u_star = np.asarray(fe.create_solver(p, bc, lin, linear=True)(tp))
    u0 = np.asarray(fe.zero_like_initial_guess(p, bc))
    step = u - u0
    assert np.linalg.norm(step - 1e-8 * (u_star - u0)) <= 1e-6 * 1e-8 * np.linalg.norm(u_star - u0)
