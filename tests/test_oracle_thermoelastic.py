"""Pins of the thermal-strain oracle (oracle/thermoelastic.py) by closed
forms.  CPU."""

import numpy as np

from oracle import hex8, thermoelastic
from oracle import structure as st


def _mesh(dims=(3, 2, 2), h=(1.0, 0.5, 2.0), seed=0):
    pts, cells = st.structured_grid(dims, h)
    L = np.array(dims) * np.array(h)
    inner = np.all((pts > 1e-9) & (pts < L - 1e-9), axis=1)
    pts = pts.copy()
    pts[inner] += np.random.default_rng(seed).uniform(-0.1, 0.1, (inner.sum(), 3)) * np.array(h)
    return pts, cells, L


def test_free_expansion_is_stress_free():
    pts, cells, _ = _mesh()
    m = hex8.Model(pts, cells)
    mat = hex8.Material("linear_elastic", nu=0.3)
    alpha, dT, t_ref = 1.2e-5, 80.0, 20.0
    u = (alpha * dT * pts).reshape(-1)
    E = np.linspace(50.0, 90.0, m.E)
    r = thermoelastic.residual(m, u, E, mat, np.full(m.N, t_ref + dT), alpha, t_ref)
    scale = np.abs(thermoelastic.thermal_residual(m, np.full(m.N, t_ref + dT), E, mat, alpha, t_ref)).max()
    assert np.abs(r).max() <= 1e-12 * scale


def test_clamped_block_reaction_is_three_k_alpha_dT():
    pts, cells, L = _mesh()
    m = hex8.Model(pts, cells)
    nu, E, alpha, dT = 0.25, 70.0, 2e-5, 50.0
    mat = hex8.Material("linear_elastic", nu=nu)
    r = thermoelastic.thermal_residual(m, np.full(m.N, dT), np.full(m.E, E), mat, alpha).reshape(-1, 3)
    on_x = np.isclose(pts[:, 0], L[0])
    p = E * alpha * dT / (1 - 2 * nu)
    assert abs(r[on_x, 0].sum() + p * L[1] * L[2]) <= 1e-12 * p * L[1] * L[2]
    interior = np.all((pts > 1e-9) & (pts < L - 1e-9), axis=1)
    assert np.abs(r[interior]).max() <= 1e-12 * np.abs(r).max()
