"""nla_rectrxm_noalloc (the entry point that allocates nothing) against
nla_rectrxm through the C ABI: the same kernels on the same inputs give the
same bits, on two calls in a row (the second smaller, so it runs on what the
first left in the shared scratch), for TRSM and TRMM, both sides, FP64 and
FP32; an order beyond the scratch is rejected before any device work."""
import ctypes

import pytest
import torch

pytestmark = pytest.mark.gpu
DEV = "cuda"
DT = {torch.float64: 1, torch.float32: 0}


def colmajor(rows, cols, dt, g):
    return torch.empty((cols, rows), dtype=dt, device=DEV).t().normal_(generator=g)


def tri(n, uplo, dt, g):
    a = colmajor(n, n, dt, g).mul_(n ** -0.5)
    a = torch.tril(a) if uplo == "L" else torch.triu(a)
    a = a.t().contiguous().t()
    a.diagonal().add_(2.0)
    return a


def call(fn, side, uplo, trans, func, A, B, alpha):
    n = A.shape[0]
    m = B.shape[1] if side == "L" else B.shape[0]
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    return fn(side.encode(), uplo.encode(), trans.encode(), func.encode(), DT[A.dtype], float(alpha), n, m,
              A.data_ptr(), A.stride(1), B.data_ptr(), B.stride(1), stream)


@pytest.mark.parametrize("dt", [torch.float64, torch.float32], ids=["f64", "f32"])
@pytest.mark.parametrize("side,uplo,trans,func", [("L", "L", "N", "S"), ("R", "U", "T", "S"), ("L", "U", "T", "M")])
def test_noalloc_equals_rectrxm(nla, dt, side, uplo, trans, func):
    from nextla_amd import _lib
    lib = _lib.load()
    g = torch.Generator(device=DEV)
    g.manual_seed(17)
    for n, m in ((3000, 200), (700, 45)):  # second call smaller
        A = tri(n, uplo, dt, g)
        B0 = colmajor(n, m, dt, g) if side == "L" else colmajor(m, n, dt, g)
        B1, B2 = B0.clone(), B0.clone()
        assert call(lib.nla_rectrxm, side, uplo, trans, func, A, B1, 1.5) == 0
        assert call(lib.nla_rectrxm_noalloc, side, uplo, trans, func, A, B2, 1.5) == 0
        torch.cuda.synchronize()
        assert torch.equal(B1, B2), (n, m)
        assert not torch.equal(B2, B0)


def test_noalloc_rejects_order_beyond_scratch(nla):
    from nextla_amd import _lib
    lib = _lib.load()
    g = torch.Generator(device=DEV)
    g.manual_seed(3)
    n = 40000  # FP64 TRSM: 625 inverse blocks of 32 KiB, above the scratch
    A = torch.zeros((1, 1), dtype=torch.float64, device=DEV)  # never read: rejected first
    B = colmajor(64, 8, torch.float64, g)
    B0 = B.clone()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    st = lib.nla_rectrxm_noalloc(b"L", b"L", b"N", b"S", 1, 1.0, n, 1, A.data_ptr(), n, B.data_ptr(), n, stream)
    assert st == _lib.NLA_ERR_WORKSPACE
    assert b"scratch" in lib.nla_last_error()
    torch.cuda.synchronize()
    assert torch.equal(B, B0)
    # TRMM needs no scratch at that order: accepted by the check (not run here)
    assert lib.nla_rectrxm_noalloc(b"L", b"L", b"N", b"Q", 1, 1.0, n, 1, A.data_ptr(), n, B.data_ptr(), n,
                                   stream) == -4
