"""The committed golden fixtures (tests/golden/, written by make_golden.py
from the oracle) against the oracle and the native host structure."""

import hashlib
import os

import numpy as np

from conftest import cantilever

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return np.load(os.path.join(G, name), allow_pickle=False)


def test_fixture_2x2x2_pattern_and_values():
    from oracle import hex8, krylov
    f = _load("fixture_2x2x2.npz")
    assert len(f["indices"]) == 3087 and len(f["indptr"]) == 82          # feax KAT: nse 3087, n = 81
    p, bc, _ = cantilever((2, 2, 2), (1.0, 1.0, 1.0), 2.0)
    np.testing.assert_array_equal(p.csr_indptr, f["indptr"])
    np.testing.assert_array_equal(p.csr_indices, f["indices"])
    np.testing.assert_array_equal(bc.bc_rows, f["bc_rows"])
    m = hex8.Model(p.points, p.cells.astype(np.int64), p.boundary_inds_list)
    mat = hex8.Material("linear_elastic", nu=0.3)
    tr = [hex8.constant_traction(p.boundary_inds_list[0], 1e-3)]
    u, info = krylov.linear_solve(m, mat, np.full(m.N, 70e3), tr, bc.bc_rows, bc.bc_vals, tol=1e-14, atol=0.0)
    np.testing.assert_allclose(info["data_bc"], f["data_bc"], rtol=0, atol=1e-12 * np.abs(f["data_bc"]).max())
    np.testing.assert_allclose(u, f["u"], rtol=1e-10, atol=1e-14)


def test_c1_pattern_hash_matches_fixture():
    f = _load("cantilever_50x5x5_samples.npz")
    p, _, _ = cantilever((50, 5, 5), (2.0, 2.0, 2.0), 100.0)
    assert p.csr_nse == int(f["nse"]) == 347904                       # nnz of feax's pattern, 50x5x5 box
    h1 = hashlib.sha256(p.csr_indptr.astype(np.int32).tobytes()).digest()
    h2 = hashlib.sha256(p.csr_indices.astype(np.int32).tobytes()).digest()
    assert h1 == bytes(f["indptr_sha256"]) and h2 == bytes(f["indices_sha256"])


def test_make_ke_3d_fixture_is_the_closed_form():
    """The unit-cube element stiffness fixture equals feax's make_KE_3d
    closed form (solvers/cmg.py:45-74) restated in test_oracle."""
    from oracle import structure as st
    from test_oracle import make_KE_3d
    K = _load("make_ke_3d.npz")["K"]
    perm = st.HEX8_OFFSETS @ np.array([4, 2, 1])     # feax local node -> cmg corner id
    P = (3 * perm[:, None] + np.arange(3)).reshape(-1)
    np.testing.assert_allclose(K, make_KE_3d()[np.ix_(P, P)], rtol=0, atol=1e-15)


def test_simp_fixture_sensitivity_and_compliance():
    f = _load("simp_4x2x2.npz")
    rho, u, dc = f["rho"], f["u"], f["dc"]
    # self-adjoint identity: sum_e rho_e dC/drho_e = -p * sum_e E'... = -3 C when Emin -> 0
    assert abs((rho * dc).sum() + 3.0 * float(f["compliance"])) <= 1e-6 * float(f["compliance"])
    assert np.all(dc < 0)
