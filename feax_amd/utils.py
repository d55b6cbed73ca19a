"""zero_like_initial_guess (feax/utils.py:464-499)."""

import numpy as np


def zero_like_initial_guess(problem, bc, device=None):
    """Zeros with the Dirichlet values set at bc_rows.

    Returns a numpy array (device=None) or a device tensor (device given).
    """
    u = np.zeros(problem.num_total_dofs_all_vars)
    u[np.asarray(bc.bc_rows, np.int64)] = bc.vals_numpy()
    if device is None:
        return u
    import torch
    return torch.as_tensor(u, device=device)
