"""Host tables of the upper-block storage of a symmetric block-3 pattern
(feax_amd.assembler.upper_block_tables): every block of the CSR pattern must
find its values through the code word -- its own stored block (column >= row
node) or the transpose of its mirror (column < row node) -- and the
workgroup column list.  Checked against a numpy SpMV that follows exactly
those addressing rules and scipy's product of the full matrix."""

import numpy as np
import pytest
import scipy.sparse as sp


def _block_pattern(problem):
    ip = np.asarray(problem.csr_indptr, np.int64)
    ix = np.asarray(problem.csr_indices, np.int64)
    N = problem.num_total_nodes
    r0 = ip[3 * np.arange(N)]
    deg = (ip[3 * np.arange(N) + 1] - r0) // 3
    first = np.repeat(r0, deg) + 3 * (np.arange(deg.sum()) - np.repeat(np.cumsum(deg) - deg, deg))
    return ip, ix, ix[first] // 3, deg


def _hex8_problem(dims):
    import feax_amd as fe

    class LE(fe.Problem):
        def get_tensor_map(self):
            return fe.materials.IsotropicLinearElastic(nu=0.3)

    mesh = fe.StructuredGrid(dims, spacing=(1.0, 1.0, 1.0)).to_mesh()
    return LE(mesh, vec=3, dim=3)


def _sym_spmv(tab, indptr, bcol, upper, x, G):
    """y = A x reading the values only through the upper-block tables, the way
    the symmetric SpMV kernel does (workgroup column lists, code words)."""
    N = (len(indptr) - 1) // 3
    y = np.zeros(3 * N)
    code = tab["code"].astype(np.int64)
    for g in range(N):
        b = g // G
        ucol = tab["ucol"][tab["uptr"][b]:tab["uptr"][b + 1]]
        k0 = indptr[3 * g] // 9
        nb = (indptr[3 * g + 1] - indptr[3 * g]) // 3
        assert np.all(code[32 * g + nb:32 * (g + 1)] == 0xFFFF)   # past the degree
        for kb in range(nb):
            w = code[32 * g + kb]
            l, low, r = w >> 6, (w >> 5) & 1, w & 31
            c = ucol[l]
            assert c == bcol[k0 + kb]
            assert bool(low) == (c < g)
            owner = c if low else g
            assert tab["usup"][tab["uptr"][b] + l] == tab["up_ptr"][c]
            blk = upper[tab["up_ptr"][owner] + r].reshape(3, 3)
            K = blk.T if low else blk
            y[3 * g:3 * g + 3] += K @ x[3 * c:3 * c + 3]
    return y


@pytest.mark.parametrize("dims,G", [((4, 3, 2), 64), ((5, 4, 3), 8)])
def test_upper_block_tables_reproduce_the_full_product(dims, G):
    from feax_amd.assembler import upper_block_tables
    p = _hex8_problem(dims)
    ip, ix, bcol, deg = _block_pattern(p)
    tab = upper_block_tables(ip, bcol, G)
    assert tab is not None and tab["nodes_per_block"] == G and tab["max_degree"] == deg.max()
    N = p.num_total_nodes
    rng = np.random.default_rng(7)
    # a random symmetric matrix on the pattern: K_cn = K_nc^T for every block pair
    node = np.repeat(np.arange(N), deg)
    keys = node * N + bcol
    mirror = np.searchsorted(keys, bcol * N + node)
    vals = rng.standard_normal((len(bcol), 3, 3))
    lower = bcol < node
    vals[lower] = vals[mirror[lower]].transpose(0, 2, 1)
    diag = bcol == node
    vals[diag] = vals[diag] + vals[diag].transpose(0, 2, 1)
    upper = vals[~lower].reshape(-1, 9)
    assert upper.shape[0] == tab["num_upper"] == tab["up_ptr"][-1]
    # the full matrix in feax's CSR layout (three rows per node, blocks in column order)
    data = np.zeros(len(ix))
    for g in range(N):
        r0, k0, d = ip[3 * g], ip[3 * g] // 9, deg[g]
        for i in range(3):
            data[r0 + i * 3 * d:r0 + (i + 1) * 3 * d] = vals[k0:k0 + d, i, :].reshape(-1)
    A = sp.csr_matrix((data, ix, ip), shape=(3 * N, 3 * N))
    assert abs(A - A.T).max() == 0
    x = rng.standard_normal(3 * N)
    y = _sym_spmv(tab, ip, bcol, upper, x, G)
    assert np.allclose(y, A @ x, rtol=1e-13, atol=1e-13)
    # the code's column position is the workgroup's sorted unique column list
    for b in range(len(tab["uptr"]) - 1):
        u = tab["ucol"][tab["uptr"][b]:tab["uptr"][b + 1]]
        assert np.all(np.diff(u) > 0)
    assert tab["max_unique"] == np.diff(tab["uptr"]).max()


def test_upper_block_tables_refuse_what_the_kernels_cannot_take():
    import feax_amd as fe
    from feax_amd.assembler import upper_block_tables
    from feax_amd.mesh import hex20_mesh
    p = _hex8_problem((3, 2, 2))
    ip, ix, bcol, deg = _block_pattern(p)
    assert upper_block_tables(ip, bcol) is not None
    # a non-symmetric pattern: node 0's last neighbour replaced by a node that
    # does not list node 0 (columns stay sorted)
    N = p.num_total_nodes
    bad = bcol.copy()
    k = deg[0] - 1
    nb0 = set(bcol[:deg[0]].tolist())
    cand = [c for c in range(bcol[k] + 1, N) if c not in nb0]
    assert cand
    bad[k] = cand[0]
    assert upper_block_tables(ip, bad) is None
    # a node without its diagonal block
    nodiag = bcol.copy()
    nodiag[0] = bcol[1] - 1 if bcol[1] - 1 > 0 else bcol[0]
    if nodiag[0] != bcol[0]:
        assert upper_block_tables(ip, nodiag) is None

    # degree above 32: a HEX20 pattern (up to 81 neighbours)
    class LE20(fe.Problem):
        def get_tensor_map(self):
            return fe.materials.IsotropicLinearElastic(nu=0.3)

    q = LE20(hex20_mesh((2, 2, 2)), vec=3, dim=3, ele_type="HEX20")
    ip2, _, bcol2, deg2 = _block_pattern(q)
    assert deg2.max() > 32 and upper_block_tables(ip2, bcol2) is None
    # at most 1023 unique column nodes per workgroup
    big = _hex8_problem((12, 12, 12))
    ipb, _, bcolb, _ = _block_pattern(big)
    assert upper_block_tables(ipb, bcolb, nodes_per_block=1) is not None
    assert upper_block_tables(ipb, bcolb, nodes_per_block=big.num_total_nodes) is None


def _star_pattern(N):
    """Block-3 pattern of N nodes: node N-1 is coupled to every node, the
    others only to themselves and to it (block columns sorted)."""
    cols = [[n, N - 1] for n in range(N - 1)] + [list(range(N))]
    deg = np.array([len(c) for c in cols])
    ip = np.zeros(3 * N + 1, np.int64)
    start = np.concatenate([[0], np.cumsum(9 * deg)])
    for n in range(N):
        for i in range(3):
            ip[3 * n + i] = start[n] + 3 * deg[n] * i
    ip[-1] = start[-1]
    return ip, np.concatenate([np.array(c) for c in cols])


def test_upper_block_tables_code_slots_of_a_star_pattern():
    """A node with 31 lower neighbours fills its 32 code slots; 33 blocks
    are refused."""
    from feax_amd.assembler import upper_block_tables
    ip, bcol = _star_pattern(32)            # node 31: 31 lower blocks + its own
    tab = upper_block_tables(ip, bcol)
    assert tab is not None and tab["max_degree"] == 32
    assert np.all(tab["code"][32 * 31:32 * 32] != 0xFFFF)
    assert np.all((tab["code"][32 * 31:32 * 31 + 31].astype(np.int64) >> 5) & 1 == 1)   # lower blocks
    assert np.all(tab["code"][2:32] == 0xFFFF)                                          # node 0: degree 2
    ip, bcol = _star_pattern(33)
    assert upper_block_tables(ip, bcol) is None


def test_morton_order_is_a_permutation_along_the_z_curve():
    from feax_amd.assembler import morton_order
    import feax_amd as fe
    pts = fe.StructuredGrid((3, 3, 3), spacing=(1.0, 1.0, 1.0)).to_mesh().points
    o = morton_order(pts)
    assert sorted(o.tolist()) == list(range(len(pts)))
    # the first 8 nodes on the curve are the unit cube at the origin
    assert {tuple(p) for p in pts[o[:8]]} == {(x, y, z) for x in (0., 1.) for y in (0., 1.) for z in (0., 1.)}
    q = np.array([[0.0, 0.0], [1.0, 1.0], [1.0, 0.0], [0.0, 1.0]])
    assert morton_order(q).tolist() == [0, 3, 2, 1]
