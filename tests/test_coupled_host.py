"""Host structure of the monolithic (u, T) problem (Problem(mesh=[m, m],
vec=[3, 1])): DOF layout and offsets, the native pattern bit-equal to the
oracle's argsort structure over the 32 x 32 element blocks, the element
slot map, and per-variable Dirichlet rows (variable_index)."""

import numpy as np
import pytest

from oracle import coupled, structure as st


class _Coupled:
    @staticmethod
    def problem(dims=(3, 2, 2), **kw):
        import feax_amd as fe

        class UT(fe.Problem):
            def get_weak_form(self):
                return fe.materials.ThermoElasticCoupled(nu=0.3, alpha=1e-2, t_ref=0.0, k=2.0)

        mesh = fe.StructuredGrid(dims).to_mesh()
        return UT(mesh=[mesh, mesh], vec=[3, 1], dim=3, ele_type=["HEX8", "HEX8"], **kw)


def test_layout_pattern_and_slots_match_oracle():
    p = _Coupled.problem(location_fns=[lambda x: np.isclose(x[0], 3.0)])
    N = p.num_total_nodes
    assert p.coupled and p.num_vars == 2 and p.offset == [0, 3 * N] and p.num_total_dofs_all_vars == 4 * N
    assert p.kernel_tag == "hex8ut" and p.solution_layout == ((N, 3), (N, 1))
    m = coupled.Model(p.points, p.cells, p.boundary_inds_list)
    assert p.csr_nse == m.nse
    assert np.array_equal(p.csr_indptr, m.csr["indptr"]) and np.array_equal(p.csr_indices, m.csr["indices"])
    assert np.array_equal(p.csr_volume_slots, m.csr["volume_slots"])
    assert np.array_equal(p._element_inds(p.cells), m.inds)
    u, T = p.unflatten_fn_sol_list(np.arange(4 * N, dtype=float))
    assert u.shape == (N, 3) and T.shape == (N, 1) and T[0, 0] == 3 * N


def test_dirichlet_rows_per_variable():
    import feax_amd as fe
    p = _Coupled.problem()
    N = p.num_total_nodes
    left = lambda x: np.isclose(x[0], 0.0)
    bottom = lambda x: np.isclose(x[2], 0.0)
    specs = [fe.DirichletBCSpec(left, "all", 0.0, variable_index=0),
             fe.DirichletBCSpec(left, 0, 1.5, variable_index=1),
             fe.DirichletBCSpec(bottom, 0, lambda x: 0.2 * x[0], variable_index=1),
             fe.DirichletBCSpec(bottom, "z", 0.0)]                      # no variable_index: every variable
    bc = fe.DirichletBCConfig(specs).create_bc(p)
    rows, vals = coupled.bc_rows(p.points, [(left, 0, "all", 0.0), (left, 1, 0, 1.5),
                                            (bottom, 1, 0, lambda x: 0.2 * x[0]), (bottom, 0, 2, 0.0)])
    # the unnamed "z" spec also hits variable 1?  component 2 >= vec 1: skipped there (feax/DCboundary.py:447-452)
    assert np.array_equal(bc.bc_rows, rows) and np.allclose(bc.bc_vals, vals)
    assert bc.bc_mask.sum() == len(rows) and bc.total_dofs == 4 * N


def test_coupled_problem_validation():
    import feax_amd as fe
    mesh = fe.StructuredGrid((2, 2, 2)).to_mesh()
    with pytest.raises(NotImplementedError):
        fe.Problem(mesh=[mesh, mesh], vec=[3, 3], dim=3)
    with pytest.raises(NotImplementedError):
        fe.Problem(mesh=[mesh, mesh, mesh], vec=[3, 1, 1], dim=3)
    other = fe.StructuredGrid((2, 2, 3)).to_mesh()
    with pytest.raises(NotImplementedError):
        fe.Problem(mesh=[mesh, other], vec=[3, 1], dim=3)
