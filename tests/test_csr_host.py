"""Host-side CSR helpers that need no GPU kernels: the pattern and gather map
of expand_view (the full matrix of an UPPER / LOWER view), run on CPU
tensors."""

import numpy as np
import pytest
import scipy.sparse as sp


@pytest.mark.parametrize("view", ["UPPER", "LOWER"])
def test_expand_view_is_a_plus_at_minus_diag(view):
    import torch
    from feax_amd.csr import CSRMatrix, expand_view
    rng = np.random.default_rng(1)
    n = 40
    M = sp.random(n, n, density=0.15, random_state=3, format="csr")
    S = (M + M.T + sp.identity(n) * 5).tocsr()
    S.sort_indices()
    T = (sp.triu(S) if view == "UPPER" else sp.tril(S)).tocsr()
    T.sort_indices()
    A = CSRMatrix(torch.as_tensor(T.data), torch.as_tensor(T.indptr.astype(np.int32)),
                  torch.as_tensor(T.indices.astype(np.int32)), (n, n))
    F = expand_view(A, view)
    got = sp.csr_matrix((F.data.numpy(), F.indices.numpy(), F.indptr.numpy()), shape=(n, n))
    assert got.has_sorted_indices
    assert np.array_equal(got.indptr, S.indptr) and np.array_equal(got.indices, S.indices)
    assert np.array_equal(got.data, S.data)
    assert expand_view(A, "FULL") is A and expand_view(A, None) is A
    wrong = "LOWER" if view == "UPPER" else "UPPER"
    with pytest.raises(ValueError, match="outside that triangle"):
        expand_view(A, wrong)
    x = rng.standard_normal(n)
    assert np.allclose(got @ x, T @ x + T.T @ x - T.diagonal() * x, rtol=1e-14, atol=1e-14)


def test_transpose_maps_match_scipy():
    import torch
    from feax_amd.csr import CSRMatrix, _transpose_maps
    M = sp.random(30, 30, density=0.2, random_state=5, format="csr")
    M.sort_indices()
    A = CSRMatrix(torch.as_tensor(M.data), torch.as_tensor(M.indptr.astype(np.int32)),
                  torch.as_tensor(M.indices.astype(np.int32)), M.shape)
    ip, ix, perm = _transpose_maps(A)
    T = M.T.tocsr()
    T.sort_indices()
    assert np.array_equal(ip.numpy(), T.indptr) and np.array_equal(ix.numpy(), T.indices)
    assert np.array_equal(A.data[perm].numpy(), T.data)
