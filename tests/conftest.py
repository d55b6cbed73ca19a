import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def pytest_collection_modifyitems(config, items):
    if gpu_available():
        return
    skip = pytest.mark.skip(reason="no GPU visible")
    for it in items:
        if "gpu" in it.keywords:
            it.add_marker(skip)


# ---------------------------------------------------------------- problems
E_MOD, NU, TRACTION = 70e3, 0.3, 1e-3


def cantilever(dims, spacing, length, material="node"):
    """feax_amd Problem + BC + TracedParams of the benchmark cantilever:
    clamped at x=0, traction t e_z on x=length (examples/benchmark/
    bench_linear_elasticity.py:323-361, surface map [0, 0, t])."""
    import feax_amd as fe

    class LinearElasticity(fe.Problem):
        def get_tensor_map(self):
            return fe.materials.IsotropicLinearElastic(nu=NU)

        def get_surface_maps(self):
            return [fe.materials.SurfaceTraction((0.0, 0.0, 1.0))]

    mesh = fe.StructuredGrid(dims, spacing=spacing).to_mesh()
    problem = LinearElasticity(mesh, vec=3, dim=3, location_fns=[lambda p: np.isclose(p[0], length)])
    bc = fe.DirichletBCConfig([fe.DirichletBCSpec(lambda p: np.isclose(p[0], 0.0), "all", 0.0)]).create_bc(problem)
    if material == "node":
        E = fe.TracedParams.create_node_var(problem, E_MOD)
    else:
        E = fe.TracedParams.create_cell_var(problem, E_MOD)
    t = fe.TracedParams.create_uniform_surface_var(problem, TRACTION)
    tp = fe.TracedParams(volume_vars=(E,), surface_vars=[(t,)])
    return problem, bc, tp


def oracle_model(problem):
    from oracle import hex8
    return hex8.Model(problem.points, problem.cells.astype(np.int64), problem.boundary_inds_list)
