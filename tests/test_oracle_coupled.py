"""Pins of the monolithic (u, T) oracle (oracle/coupled.py)."""

import numpy as np
import scipy.sparse.linalg as spla

from oracle import conduction, coupled, hex8, krylov, structure as st, thermoelastic

LAW = (0.3, 1.2e-2, 0.5, 2.5)       # nu, alpha, t_ref, k


def _model(dims=(4, 2, 2), distort=0.1, seed=0, load=True):
    pts, cells = st.structured_grid(dims)
    if distort:
        rng = np.random.default_rng(seed)
        inner = np.all((pts > 1e-9) & (pts < np.array(dims) - 1e-9), axis=1)
        pts = pts + rng.uniform(-distort, distort, pts.shape) * inner[:, None]
    faces = [st.boundary_faces(pts, cells, lambda p: np.isclose(p[0], dims[0])),
             st.boundary_faces(pts, cells, lambda p: np.isclose(p[2], dims[2]))] if load else []
    return coupled.Model(pts, cells, faces), pts, cells, faces


def test_jacobian_is_the_residual_derivative_and_blocks_match():
    m, pts, cells, faces = _model()
    rng = np.random.default_rng(1)
    E = rng.uniform(1.0, 2.0, m.N)
    K = m.scipy_csr(m.csr_values(E, LAW)).toarray()
    w = rng.standard_normal(m.n)
    v = rng.standard_normal(m.n)
    fd = (m.residual(w + v, E, LAW) - m.residual(w - v, E, LAW)) / 2.0      # exact: R is affine in w
    assert np.allclose(K @ v, fd, atol=1e-12 * np.abs(fd).max())
    N = m.N
    uu, ut, tu, tt = K[:3 * N, :3 * N], K[:3 * N, 3 * N:], K[3 * N:, :3 * N], K[3 * N:, 3 * N:]
    h = hex8.Model(pts, cells)
    Kh = h.scipy_csr(h.csr_values(np.zeros(h.n), E, hex8.Material("linear_elastic", nu=LAW[0]))).toarray()
    assert np.allclose(uu, Kh, atol=1e-14 * np.abs(Kh).max())
    c = conduction.Model(pts, cells)
    Kc = c.scipy_csr(c.csr_values(np.full(m.E, LAW[3]))).toarray()
    assert np.allclose(tt, Kc, atol=1e-14 * np.abs(Kc).max())
    assert np.abs(tu).max() == 0.0 and np.abs(ut).max() > 0
    # K_uT T = the thermal residual of oracle/thermoelastic (with t_ref = 0)
    T = rng.standard_normal(N)
    Rth = thermoelastic.thermal_residual(h, T, E, hex8.Material("linear_elastic", nu=LAW[0]), LAW[1], 0.0)
    assert np.allclose(ut @ T, Rth, atol=1e-13 * np.abs(Rth).max())
    # the pattern: every (u|T row, u|T column) pair of neighbouring nodes
    assert m.nse == 16 * len(h.csr["indices"]) // 9


def test_monolithic_solution_equals_staggered():
    dims = (5, 2, 3)
    m, pts, cells, faces = _model(dims, distort=0.08, seed=2)
    N = m.N
    rng = np.random.default_rng(3)
    E = rng.uniform(1.0, 2.0, m.E)                                           # cell variable
    tz = rng.uniform(0.1, 0.3, (len(faces[0]), 4))
    qz = rng.uniform(0.5, 1.0, (len(faces[1]), 4))
    loads = [(0, (0.0, 0.0, 1.0), tz), (1, None, qz)]
    specs = [(lambda p: np.isclose(p[0], 0.0), 0, "all", 0.0),
             (lambda p: np.isclose(p[0], 0.0), 1, 0, 1.5),
             (lambda p: np.isclose(p[2], 0.0), 1, 0, lambda p: 0.2 * p[0])]
    rows, vals = coupled.bc_rows(pts, specs)
    mask = np.zeros(m.n, bool)
    mask[rows] = True
    w0 = np.zeros(m.n)
    w0[rows] = vals
    K = m.scipy_csr(hex8.apply_bc_csr(m.csr_values(E, LAW), m.csr, mask, rows, True))
    b = -hex8.apply_bc_res(m.residual(w0, E, LAW, loads), w0, rows, vals)
    w = w0 + spla.spsolve(K.tocsc(), b)
    # staggered: conduction for T, then the thermo-elastic solve with that T
    c = conduction.Model(pts, cells, [faces[1]])
    trow = rows[rows >= 3 * N] - 3 * N
    tval = vals[rows >= 3 * N]
    tmask = np.zeros(N, bool)
    tmask[trow] = True
    T0 = np.zeros(N)
    T0[trow] = tval
    Kc = c.scipy_csr(hex8.apply_bc_csr(c.csr_values(np.full(m.E, LAW[3])), c.csr, tmask, trow, True))
    bc_ = -hex8.apply_bc_res(c.residual(T0, np.full(m.E, LAW[3]), fluxes=[qz]), T0, trow, tval)
    T = T0 + spla.spsolve(Kc.tocsc(), bc_)
    h = hex8.Model(pts, cells, [faces[0]])
    mat = hex8.Material("linear_elastic", nu=LAW[0])
    urow = rows[rows < 3 * N]
    uval = vals[rows < 3 * N]
    umask = np.zeros(3 * N, bool)
    umask[urow] = True
    u0 = np.zeros(3 * N)
    u0[urow] = uval
    Kh = h.scipy_csr(hex8.apply_bc_csr(h.csr_values(u0, E, mat), h.csr, umask, urow, True))
    trac = [tz[:, :, None] * np.array([0.0, 0.0, 1.0])]
    Rh = thermoelastic.residual(h, u0, E, mat, T, LAW[1], LAW[2], trac)
    u = u0 + spla.spsolve(Kh.tocsc(), -hex8.apply_bc_res(Rh, u0, urow, uval))
    assert np.allclose(w[3 * N:], T, rtol=1e-11, atol=1e-12 * np.abs(T).max())
    assert np.allclose(w[:3 * N], u, rtol=1e-10, atol=1e-11 * np.abs(u).max())
    assert np.abs(u).max() > 1e-3 and np.abs(T).max() > 1.0
