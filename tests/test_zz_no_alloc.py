"""fx_pcg_f64_noalloc (the PCG whose scratch is a static device array)
against fx_pcg_f64 (caller workspace) through the C-ABI: two calls in a row,
the second on a smaller system, each equal to the workspace entry point's
result on the same inputs (same kernels in the same order: bit-equal
iterates and iteration counts), and a size beyond the array refused."""

import ctypes

import numpy as np
import pytest

from conftest import cantilever

pytestmark = pytest.mark.gpu


def _system(dims, length):
    import torch
    import feax_amd as fe
    from feax_amd.assembler import assemble_csr, residual
    p, bc, tp = cantilever(dims, (1.0, 1.0, 1.0), length)
    ts = fe.TracedStructure.from_problem(p)
    A, dinv = assemble_csr(p, ts, tp, None, bc, True, torch.float64, jacobi_shift=1e-12)
    b = residual(p, ts, tp, torch.zeros(ts.n, dtype=torch.float64, device="cuda"), bc, negate=True)
    return A, dinv, b


def _solve(name, A, dinv, b, flags, maxiter=20000):
    import torch
    from feax_amd import _native as nat
    n = A.shape[0]
    x = torch.full((n,), float("nan"), dtype=torch.float64, device="cuda")
    iters = ctypes.c_int32(-1)
    rs = ctypes.c_double(-1.0)
    bcol = A.bcol if flags & nat.FX_PCG_BLOCK3 else None
    args = [nat.stream(), n, A.nse, flags, nat.ptr(A.indptr), nat.ptr(A.indices), nat.ptr(bcol), nat.ptr(A.data),
            nat.ptr(dinv), nat.ptr(b), None, 1e-10, 0.0, maxiter, 32, nat.ptr(x), ctypes.addressof(iters),
            ctypes.addressof(rs)]
    if name == "fx_pcg_f64":
        ws = torch.empty(int(nat.lib().fx_pcg_workspace_bytes(n, 8)), dtype=torch.uint8, device="cuda")
        args += [None, None, 0, nat.ptr(ws), ws.numel()]
    nat.call(name, *args)
    torch.cuda.synchronize()
    return x.cpu().numpy(), iters.value, rs.value


@pytest.mark.parametrize("block3", [True, False])
def test_noalloc_pcg_matches_workspace_pcg_twice(block3):
    from feax_amd import _native as nat
    flags = nat.FX_PCG_BLOCK3 if block3 else 0
    for dims in ((40, 6, 6), (12, 3, 3)):       # second call: a smaller system
        print("case: fx_pcg_f64_noalloc vs fx_pcg_f64", dims, "block3" if block3 else "csr", flush=True)
        A, dinv, b = _system(dims, float(dims[0]))
        x_ref, it_ref, rs_ref = _solve("fx_pcg_f64", A, dinv, b, flags)
        x, it, rs = _solve("fx_pcg_f64_noalloc", A, dinv, b, flags)
        assert it_ref > 10 and it == it_ref and rs == rs_ref
        assert np.array_equal(x, x_ref)
        bn = np.linalg.norm(b.cpu().numpy())
        assert rs ** 0.5 <= 1e-10 * bn


def test_noalloc_pcg_refuses_sizes_beyond_its_scratch():
    import torch
    from feax_amd import _native as nat
    n = 1231169 * 3         # beyond the 40 MiB array; nothing is launched
    x = torch.empty(3, dtype=torch.float64, device="cuda")
    iters = ctypes.c_int32(0)
    rs = ctypes.c_double(0.0)
    rc = nat.lib().fx_pcg_f64_noalloc(nat.stream(), n, 9, 0, nat.ptr(x), nat.ptr(x), None, nat.ptr(x), None,
                                      nat.ptr(x), None, 1e-8, 0.0, 10, 32, nat.ptr(x), ctypes.addressof(iters),
                                      ctypes.addressof(rs))
    assert rc == 1002           # FX_ELIMIT
    assert "scratch" in nat.last_error()
