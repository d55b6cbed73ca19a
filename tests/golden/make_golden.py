"""Writes the golden fixtures of tests/golden/ from the FP64 oracle (oracle/).

    python tests/golden/make_golden.py            # small fixtures (seconds)
    python tests/golden/make_golden.py --large    # + 200x40x40 and 120x40x20 oracle solves (CPU, minutes)

Fixtures (npz, data only):
  fixture_2x2x2.npz   the 2x2x2 unit cantilever of feax's conftest (clamped
                      x=0, traction 1e-3 e_z on x=2, E=70e3 node var, nu=0.3):
                      indptr, indices, BC-eliminated CSR data, residual at
                      u=0, b, direct-solve u
  cantilever_50x5x5_samples.npz      50x5x5 (h=2): sha256 of indptr/indices, 1000 seeded
                      (slot, value) pairs of the BC-eliminated CSR, u at 64
                      seeded DOFs (direct solve)
  make_ke_3d.npz      unit-cube element stiffness (E=1, nu=0.3), feax corner order
  simp_4x2x2.npz      SIMP compliance sensitivity on a 4x2x2 grid (rho seeded)
  cantilever_200x40x40_oracle.npz       200x40x40 (h=1) cantilever: oracle Jacobi-PCG to rel
                      1e-12, u at 4096 seeded DOFs, max|u|, ||u||
  simp_120x40x20_oracle.npz       120x40x20 SIMP cantilever (rho ~ U(0.3,1) seeded,
                      unit patch load): oracle PCG to rel 1e-12, compliance,
                      closed-form dC/drho at 4096 seeded cells, max|dC/drho|
  batch64_oracle.npz  the seeded set of 64 uneven SIMP cantilevers (nx ~ U{20..80},
                      ny, nz ~ U{8..24}, rho ~ U(0.3,1), traction 10^U(-4,-2)):
                      dims of all 64, and for 5-6 of them (the smallest, the
                      largest, the median, the first three) the oracle Jacobi-PCG solution to
                      rel 1e-12 at 1024 seeded DOFs, max|u|
  neohookean_100x20x20_oracle.npz   Neo-Hookean block 100x20x20 (mu ~ U(0.5,1.5) per
                      cell, kappa 10, bottom clamped, top u_z = -0.4 s, s = 1..5):
                      oracle Newton (Jacobi-PCG solves to rel 1e-13, started from
                      the linearised predictor, to |R| <= 1e-11) per step; final
                      state at 4096 seeded DOFs, max|u|, Newton iterations
"""

import argparse
import hashlib
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import adjoint, hex8, krylov, structure  # noqa: E402

SEED = 20261203


def cantilever_model(dims, h, t=1e-3, E=70e3):
    pts, cells = structure.structured_grid(dims, (h, h, h))
    L = dims[0] * h
    faces = structure.boundary_faces(pts, cells, lambda p: np.isclose(p[0], L))
    m = hex8.Model(pts, cells, [faces])
    bc_rows, bc_vals = structure.dirichlet_rows(pts, [(lambda p: np.isclose(p[0], 0.0), "all", 0.0)])
    return m, np.full(m.N, E), [hex8.constant_traction(faces, t)], bc_rows, bc_vals


def fixture_2x2x2():
    import scipy.sparse.linalg as spla
    m, E, tr, rows, vals = cantilever_model((2, 2, 2), 1.0)
    mat = hex8.Material("linear_elastic", nu=0.3)
    u, info = krylov.linear_solve(m, mat, E, tr, rows, vals, tol=1e-14, atol=0.0)
    ud = spla.spsolve(info["K"].tocsc(), info["b"])
    res0 = m.residual(np.zeros(m.n), E, mat, tr)
    np.savez(os.path.join(HERE, "fixture_2x2x2.npz"), indptr=m.csr["indptr"], indices=m.csr["indices"],
             data_bc=info["data_bc"], res0=res0, b=info["b"], u=ud, bc_rows=rows)


def cantilever_50x5x5_samples():
    import scipy.sparse.linalg as spla
    m, E, tr, rows, vals = cantilever_model((50, 5, 5), 2.0)
    mat = hex8.Material("linear_elastic", nu=0.3)
    _, info = krylov.linear_solve(m, mat, E, tr, rows, vals, tol=1e-12, atol=0.0)
    ud = spla.spsolve(info["K"].tocsc(), info["b"])
    rng = np.random.default_rng(SEED)
    slots = np.sort(rng.choice(m.nse, 1000, replace=False))
    dofs = np.sort(rng.choice(m.n, 64, replace=False))
    np.savez(os.path.join(HERE, "cantilever_50x5x5_samples.npz"),
             indptr_sha256=np.frombuffer(hashlib.sha256(m.csr["indptr"].astype(np.int32).tobytes()).digest(), np.uint8),
             indices_sha256=np.frombuffer(hashlib.sha256(m.csr["indices"].astype(np.int32).tobytes()).digest(),
                                          np.uint8),
             nse=m.nse, slots=slots, values=info["data_bc"][slots], dofs=dofs, u=ud[dofs])


def make_ke_3d():
    pts, cells = structure.structured_grid((1, 1, 1))
    m = hex8.Model(pts, cells, structure=False)
    K = m.element_matrices(np.zeros(m.n), np.ones(1), hex8.Material("linear_elastic", nu=0.3))[0]
    np.savez(os.path.join(HERE, "make_ke_3d.npz"), K=K)


def simp_4x2x2():
    import scipy.sparse.linalg as spla
    pts, cells = structure.structured_grid((4, 2, 2))
    faces = structure.boundary_faces(pts, cells, lambda p: np.isclose(p[0], 4.0))
    m = hex8.Model(pts, cells, [faces])
    rows, vals = structure.dirichlet_rows(pts, [(lambda p: np.isclose(p[0], 0.0), "all", 0.0)])
    mat = hex8.Material("simp", nu=0.3, E0=1.0, Emin=1e-9, penal=3.0)
    rho = np.random.default_rng(SEED).uniform(0.3, 1.0, m.E)
    mask = np.zeros(m.n, bool)
    mask[rows] = True
    K = m.scipy_csr(hex8.apply_bc_csr(m.csr_values(np.zeros(m.n), rho, mat), m.csr, mask, rows))
    tr = [hex8.constant_traction(faces, 0.25)]
    f = -m.residual(np.zeros(m.n), rho, mat, tr)
    f[rows] = 0.0
    u = spla.spsolve(K.tocsc(), f)
    np.savez(os.path.join(HERE, "simp_4x2x2.npz"), rho=rho, u=u,
             compliance=adjoint.compliance(m, u, [faces], tr),
             dc=adjoint.compliance_sens_closed_form(m, u, rho, mat))


def cantilever_200x40x40_oracle():
    t0 = time.time()
    m, E, tr, rows, vals = cantilever_model((200, 40, 40), 1.0)
    print(f"200x40x40 model {time.time() - t0:.0f} s", flush=True)
    mat = hex8.Material("linear_elastic", nu=0.3)
    u, info = krylov.linear_solve(m, mat, E, tr, rows, vals, tol=1e-12, atol=0.0, maxiter=100000)
    print(f"200x40x40 solve {time.time() - t0:.0f} s, {info['iters']} iterations", flush=True)
    relres = np.linalg.norm(info["K"] @ (u - 0.0) - info["b"]) / np.linalg.norm(info["b"])
    dofs = np.sort(np.random.default_rng(SEED).choice(m.n, 4096, replace=False))
    np.savez(os.path.join(HERE, "cantilever_200x40x40_oracle.npz"), dofs=dofs, u=u[dofs], u_max=np.abs(u).max(),
             u_norm=np.linalg.norm(u), iters=info["iters"], relres=relres)


def simp_120x40x20_oracle():
    t0 = time.time()
    dims = (120, 40, 20)
    pts, cells = structure.structured_grid(dims)
    faces = structure.boundary_faces(
        pts, cells, lambda p: np.isclose(p[0], 120) & (np.abs(p[1] - 20) <= 1 + 1e-9) & (np.abs(p[2] - 10) <= 1 + 1e-9))
    assert len(faces) == 4
    m = hex8.Model(pts, cells, [faces])
    rows, vals = structure.dirichlet_rows(pts, [(lambda p: np.isclose(p[0], 0.0), "all", 0.0)])
    rho = np.random.default_rng(SEED).uniform(0.3, 1.0, m.E)
    mat = hex8.Material("simp", nu=0.3, E0=1.0, Emin=1e-9, penal=3.0)
    tr = [hex8.constant_traction(faces, 0.25)]
    u, info = krylov.linear_solve(m, mat, rho, tr, rows, vals, tol=1e-12, atol=0.0, maxiter=100000)
    print(f"c3 solve {time.time() - t0:.0f} s, {info['iters']} iterations", flush=True)
    dc = adjoint.compliance_sens_closed_form(m, u, rho, mat)
    cells_s = np.sort(np.random.default_rng(SEED).choice(m.E, 4096, replace=False))
    np.savez(os.path.join(HERE, "simp_120x40x20_oracle.npz"), cells=cells_s, dc=dc[cells_s], dc_max=np.abs(dc).max(),
             compliance=adjoint.compliance(m, u, [faces], tr), iters=info["iters"])


def batch64_instances(seed=SEED, nseg=64):
    """The 64-instance set in the draw order of the batched benchmark: per
    instance nx, ny, nz, then rho (cell order), then the traction."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(nseg):
        nx, ny, nz = int(rng.integers(20, 81)), int(rng.integers(8, 25)), int(rng.integers(8, 25))
        rho = rng.uniform(0.3, 1.0, nx * ny * nz)
        t = 10.0 ** rng.uniform(-4, -2)
        out.append(((nx, ny, nz), rho, t))
    return out


def batch64_oracle():
    inst = batch64_instances()
    dofs = [3 * (d[0] + 1) * (d[1] + 1) * (d[2] + 1) for d, _, _ in inst]
    pick = sorted({int(np.argmin(dofs)), int(np.argmax(dofs)), int(np.argsort(dofs)[len(dofs) // 2]), 0, 1, 2})
    mat = hex8.Material("simp", nu=0.3, E0=1.0, Emin=1e-9, penal=3.0)
    res = dict(dims=np.array([d for d, _, _ in inst]), tractions=np.array([t for _, _, t in inst]),
               picked=np.array(pick))
    for k in pick:
        t0 = time.time()
        dims, rho, t = inst[k]
        pts, cells = structure.structured_grid(dims)
        L = float(dims[0])
        faces = structure.boundary_faces(pts, cells, lambda p: np.isclose(p[0], L))
        m = hex8.Model(pts, cells, [faces])
        rows, vals = structure.dirichlet_rows(pts, [(lambda p: np.isclose(p[0], 0.0), "all", 0.0)])
        u, info = krylov.linear_solve(m, mat, rho, [hex8.constant_traction(faces, t)], rows, vals, tol=1e-12,
                                      atol=0.0, maxiter=200000)
        sel = np.sort(np.random.default_rng(SEED + k).choice(m.n, 1024, replace=False))
        res[f"dofs_{k}"] = sel
        res[f"u_{k}"] = u[sel]
        res[f"umax_{k}"] = np.abs(u).max()
        print(f"instance {k} {dims}: {m.n} DOF, {info['iters']} iterations, {time.time() - t0:.0f} s", flush=True)
    np.savez(os.path.join(HERE, "batch64_oracle.npz"), **res)


def neohookean_100x20x20_oracle():
    from oracle import newton
    dims = (100, 20, 20)
    nz = dims[2]
    pts, cells = structure.structured_grid(dims)
    m = hex8.Model(pts, cells)
    bottom, _ = structure.dirichlet_rows(pts, [(lambda p: np.isclose(p[2], 0.0), "all", 0.0)])
    rows, _ = structure.dirichlet_rows(pts, [(lambda p: np.isclose(p[2], 0.0), "all", 0.0),
                                             (lambda p: np.isclose(p[2], nz), 2, 0.0)])
    is_top = ~np.isin(rows, bottom)
    mu = np.random.default_rng(SEED).uniform(0.5, 1.5, m.E)
    mat = hex8.Material("neo_hookean", kappa=10.0)
    u = np.zeros(m.n)
    iters = []
    for s in range(1, 6):
        t0 = time.time()
        vals = np.where(is_top, -0.4 * s, 0.0)
        # The converged state does not depend on the initial guess; the
        # linearised predictor keeps every Newton tangent SPD (Jacobi-PCG
        # solves to rel 1e-13 instead of sparse LU at 134k DOF).
        u0 = newton.linearized_predictor(m, mat, mu, rows, vals, u)
        u, it, hist = newton.newton(m, mat, mu, rows, vals, u0, tol=1e-11, rel_tol=1e-12, linear="pcg")
        iters.append(it)
        print(f"step {s}: {it} Newton iterations, |R| {hist[-1]:.3e} (rel {hist[-1] / hist[0]:.2e}), "
              f"{time.time() - t0:.0f} s", flush=True)
    sel = np.sort(np.random.default_rng(SEED).choice(m.n, 4096, replace=False))
    np.savez(os.path.join(HERE, "neohookean_100x20x20_oracle.npz"), dofs=sel, u=u[sel], u_max=np.abs(u).max(),
             u_norm=np.linalg.norm(u), newton_iters=np.array(iters))


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--large", action="store_true")
    ap.add_argument("--only", default=None)
    a = ap.parse_args()
    jobs = [fixture_2x2x2, cantilever_50x5x5_samples, make_ke_3d, simp_4x2x2] + (
        [cantilever_200x40x40_oracle, simp_120x40x20_oracle, batch64_oracle, neohookean_100x20x20_oracle]
        if a.large else [])
    for job in jobs:
        if a.only and job.__name__ != a.only:
            continue
        t = time.time()
        job()
        print(f"{job.__name__}: {time.time() - t:.1f} s", flush=True)
