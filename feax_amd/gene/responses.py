"""Compliance response (feax/gene/responses.py:10-105) and the fused SIMP
compliance sensitivity (feax/solvers/cmg.py:312-323).

create_compliance_fn(problem, surface_load_params) returns compliance_fn(sol)
= sum over loaded surfaces of sum_q (-t(u, x, p)) . u nanson.  With the
dead-traction maps of this path, t = direction * p, the compliance is the
linear form C = g . u with g = -(surface residual at load p), which the
residual kernel evaluates once; compliance_fn is then a torch dot product,
differentiable in sol through torch.autograd.
"""

import ctypes

import numpy as np

from .. import _native as nat


def _load_params(problem, surface_load_params):
    num = len(problem.boundary_inds_list)
    if surface_load_params is None:
        return [1.0] * num
    if not isinstance(surface_load_params, (list, tuple, np.ndarray)):
        return [surface_load_params] * num
    if len(surface_load_params) != num:
        raise ValueError(f"surface_load_params length {len(surface_load_params)} must match number of "
                         f"surfaces {num}")
    return list(surface_load_params)


def _external_work_vector(problem, ts, params):
    """g with C = g . u: minus the traction residual of the loads `params`."""
    import torch
    from ..assembler import elem_workspace
    parts = []
    for i, p in enumerate(params):
        nf = int(ts.face_offsets[i + 1] - ts.face_offsets[i])
        v = torch.as_tensor(np.asarray(p, np.float64), device=ts.device)
        q = ts.num_face_quads
        parts.append(torch.broadcast_to(v.reshape(-1, 1) if v.numel() == nf else v, (nf, q)).to(torch.float64))
    svar = torch.cat(parts).reshape(-1).contiguous()
    # At u = 0 the volume term vanishes for every law here (stress(0) = 0).
    kind = nat.FX_MAT_CONDUCTION if ts.tag in ("hex8s", "tet4s", "quad4s") else nat.FX_MAT_LINEAR_ELASTIC
    mat = nat.material_struct(kind, nat.FX_PARAM_CELL, nu=0.3)
    ones = torch.ones(ts.num_cells, dtype=torch.float64, device=ts.device)
    zero = torch.zeros(ts.n, dtype=torch.float64, device=ts.device)
    g = torch.empty(ts.n, dtype=torch.float64, device=ts.device)
    nat.call(f"fx_{ts.tag}_residual_f64", nat.stream(), ts.num_nodes, ts.num_cells, nat.ptr(ts.x), nat.ptr(ts.y),
             nat.ptr(ts.z), nat.ptr(ts.cells), nat.ptr(ts.inc_ptr), nat.ptr(ts.inc), ctypes.addressof(mat),
             nat.ptr(ones), nat.ptr(zero), ts.num_faces, nat.ptr(ts.faces), nat.ptr(ts.face_inc_ptr),
             nat.ptr(ts.face_inc), nat.ptr(svar), nat.ptr(ts.fdir), 0, None, None, 1, nat.ptr(g),
             *elem_workspace(ts, getattr(nat.lib(), f"fx_{ts.tag}_residual_workspace_bytes")(ts.num_cells)))
    return g


def create_compliance_fn(problem, surface_load_params=None, traced_structure=None):
    """compliance_fn(sol) -> C (0-d float64 tensor; sol: flat dofs, Solution
    or torch tensor).  feax/gene/responses.py:10-105."""
    from ..assembler import _ts_or_build
    if len(problem.boundary_inds_list) == 0:
        return lambda sol: 0.0
    params = _load_params(problem, surface_load_params)
    ts = _ts_or_build(problem, traced_structure)
    g = _external_work_vector(problem, ts, params)

    def compliance_fn(sol):
        import torch
        u = getattr(sol, "dofs", sol)
        if not isinstance(u, torch.Tensor):
            u = torch.as_tensor(np.asarray(u, np.float64), device=ts.device)
        return torch.dot(g, u.reshape(-1)[: ts.n].to(torch.float64))

    compliance_fn.external_work_vector = g
    return compliance_fn


def compliance_and_sensitivity(problem, traced_params, sol, surface_load_params=None, traced_structure=None):
    """(C, dC/drho): the compliance of sol and the fused self-adjoint SIMP
    sensitivity -p rho^(p-1) (E0 - Emin) u_e^T K0_e u_e per element
    (fx_simp_compliance_sens_f64; feax/solvers/cmg.py:312-323).  Valid when
    the Dirichlet values are zero and the load does not depend on rho."""
    from ..assembler import _ts_or_build, compliance_sensitivity
    ts = _ts_or_build(problem, traced_structure)
    C = create_compliance_fn(problem, surface_load_params, ts)(sol)
    u = getattr(sol, "dofs", sol)
    return C, compliance_sensitivity(problem, ts, traced_params, u)
