"""CPU oracle: an FP64 numpy/scipy restatement of feax's assembly + Krylov path.

TEST INFRASTRUCTURE ONLY.  Nothing in ``feax_amd`` imports this package; only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may use it, and only as the checker (or, for ``cpu_baseline``, as
the timed CPU port).  The product path runs on the HIP library and fails
loudly without it.

Modules (each function cites the feax file:line it restates):

* ``structure`` -- StructuredGrid numbering, boundary faces, DOF map, raw COO
  (I, J) order and the argsort-based CSR pattern + slot maps
  (feax/spgrid.py, feax/fe.py, feax/problem.py).
* ``hex8``      -- HEX8 basis on 2x2x2 Gauss points, geometry, element
  residual / stiffness for the supported materials, surface traction,
  residual and CSR value assembly, Dirichlet elimination
  (feax/basis.py, feax/fe.py, feax/assembler.py, feax/DCboundary.py).
* ``krylov``    -- Jacobi-preconditioned CG with jax.scipy.sparse.linalg.cg
  stopping semantics, and the linear solve wrapper
  (feax/solvers/common.py, feax/solvers/linear.py).
* ``sens``      -- adjoint / SIMP compliance sensitivity and the
  closed-form unit-cube KE (feax/solvers/cmg.py, feax/gene/responses.py).
* ``newton``    -- Neo-Hookean Newton with Armijo backtracking
  (feax/solvers/newton.py).

How each part is pinned (the reference cannot be imported: jax, basix and
gmsh are absent) is listed in the module headers; functions with no pin
say "parity unpinned".
"""
