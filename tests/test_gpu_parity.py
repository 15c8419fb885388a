"""Parity of the HIP path (through the C ABI and its Python mirror) with the
oracle: OpenBLAS fixtures, the reference's own test grids, the C restatement
of unified_rectrxm!, closed-form and round-trip properties at full size.

Tolerances:
  * reference grids: relative forward error vs BLAS < 1e-14 (FP64,
    test/unified_rectrxm.jl:8) and < 1e-5 (FP32 leaves, test/trsm.jl:8);
  * everything else: normwise backward error (TRSM) / product error (TRMM)
    <= 10*n*eps of the dtype, computed in FP64.
"""
import itertools
import json
import os

import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
COMBOS = list(itertools.product("LR", "LU", "NTC", "SM"))
DEV = "cuda"


def dev(a):
    """numpy -> column-major device tensor (Julia layout)."""
    a = np.asarray(a)
    return torch.from_numpy(np.ascontiguousarray(a.T)).to(DEV).t()


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def eps(dt):
    return float(np.finfo(dt).eps)


def run(nla, side, uplo, trans, alpha, func, A, B):
    Ad, Bd = dev(A), dev(B)
    nla.unified_rectrxm_(side, uplo, trans, alpha, func, Ad, Bd)
    return host(Bd)


def check_bound(side, uplo, trans, alpha, func, A, B, X, factor=10.0):
    n = A.shape[0]
    e = (O.backward_error if func == "S" else O.product_error)(side, uplo, trans, alpha, A, B, X)
    bound = factor * max(n, 1) * eps(A.dtype)
    assert e <= bound, (side, uplo, trans, func, n, e, bound)
    return e


def test_device_is_gfx950(nla):
    assert torch.cuda.is_available()
    name = torch.cuda.get_device_properties(0).gcnArchName
    assert name.startswith("gfx950"), name


def test_blas_fixtures(nla):
    data = np.load(os.path.join(HERE, "golden", "blas_fixtures.npz"))
    cases = json.load(open(os.path.join(HERE, "golden", "blas_fixtures.json")))["cases"]
    for c in cases:
        A, B, X = data[c["A"]], data[c["B"]], data[c["key"]]
        got = run(nla, c["side"], c["uplo"], c["trans"], c["alpha"], c["func"], A, B)
        tol = 1e-14 if c["dtype"] == "float64" else 1e-5
        assert O.rel_error(got, X) < tol, (c, O.rel_error(got, X))


def test_reference_grid_fp64(nla):
    """The 288-case grid of test/unified_rectrxm.jl:11-43 (n in 16..256,
    m in {1,8,64}, all side/uplo/trans/func, alpha 1.0, FP64)."""
    rng = np.random.default_rng(0)
    for n, m in itertools.product((16, 32, 128, 256), (1, 8, 64)):
        for side, uplo, trans, func in COMBOS:
            A = np.tril(rng.random((n, n)) + 1) if uplo == "L" else np.triu(rng.random((n, n)) + 1)
            A = A + 10 * np.eye(n)
            B = rng.random((n, m) if side == "L" else (m, n)) + 1
            got = run(nla, side, uplo, trans, 1.0, func, A, B)
            ref = O.blas(side, uplo, trans, 1.0, func, A, B)
            assert O.rel_error(got, ref) < 1e-14, (n, m, side, uplo, trans, func, O.rel_error(got, ref))


def test_reference_leaf_grid_fp32(nla):
    """The 48-case FP32 leaf grid of test/trsm.jl:10-64 (LeftLowerTRSM! ...)."""
    rng = np.random.default_rng(1)
    fns = {("L", "L"): nla.LeftLowerTRSM_, ("L", "U"): nla.LeftUpperTRSM_,
           ("R", "L"): nla.RightLowerTRSM_, ("R", "U"): nla.RightUpperTRSM_}
    for (side, uplo), fn in fns.items():
        for n, m in itertools.product((16, 32, 128, 256), (1, 8, 64)):
            T = (rng.random((n, n)) + 1).astype(np.float32)
            A = (np.tril(T) if uplo == "L" else np.triu(T)) + 10 * np.eye(n, dtype=np.float32)
            B = (rng.random((n, m) if side == "L" else (m, n)) + 1).astype(np.float32)
            Ad, Bd = dev(A), dev(B)
            fn(Ad, Bd)
            ref = O.blas(side, uplo, "N", 1.0, "S", A, B)
            assert O.rel_error(host(Bd), ref) < 1e-5, (side, uplo, n, m)


def test_trmm_leaf_functions(nla):
    rng = np.random.default_rng(2)
    fns = {("L", "L"): nla.LeftLowerTRMM_, ("L", "U"): nla.LeftUpperTRMM_,
           ("R", "L"): nla.RightLowerTRMM_, ("R", "U"): nla.RightUpperTRMM_}
    for (side, uplo), fn in fns.items():
        for n, m in ((16, 5), (64, 64), (100, 3)):
            A = rng.standard_normal((n, n))
            B = rng.standard_normal((n, m) if side == "L" else (m, n))
            Ad, Bd = dev(A), dev(B)
            fn(Ad, Bd)
            check_bound(side, uplo, "N", 1.0, "M", A, B, host(Bd))


SIZES = [(1, 1), (2, 3), (63, 5), (64, 64), (65, 1), (100, 65), (129, 7), (300, 130), (513, 31), (1000, 64),
         (1031, 2)]

# methods below the recursion: fused 256-panels (default), 64-leaves as MFMA
# products of inverted blocks, 64-leaves by register/LDS substitution
METHODS = {"panel": (1, 1), "mfma-leaf": (1, 0), "subst-leaf": (0, 0)}


@pytest.fixture(params=list(METHODS))
def method(request, nla):
    from nextla_amd import _lib
    leaf_mode, panel = METHODS[request.param]
    old_l = _lib.set_config(_lib.NLA_CFG_LEAF_MODE, leaf_mode)
    old_p = _lib.set_config(_lib.NLA_CFG_PANEL, panel)
    yield request.param
    _lib.set_config(_lib.NLA_CFG_LEAF_MODE, old_l)
    _lib.set_config(_lib.NLA_CFG_PANEL, old_p)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("n,m", SIZES)
def test_recursion_all_combos(nla, method, dt, n, m):
    """All side/uplo/trans/func at sizes that hit single leaves, ragged leaves,
    odd splits and multi-level recursion; alpha != 1; backward error bound and
    agreement with the C restatement of the reference algorithm; for each
    method below the recursion."""
    rng = np.random.default_rng(n * 1000 + m)
    for side, uplo, trans, func in itertools.product("LR", "LU", "NT", "SM"):
        A = O._f(np.tril(rng.uniform(-1, 1, (n, n)) / np.sqrt(n)) + np.diag(2 + rng.random(n))).astype(dt)
        if uplo == "U":
            A = np.asfortranarray(A.T)
        B = rng.standard_normal((n, m) if side == "L" else (m, n)).astype(dt)
        alpha = 1.5 if func == "S" else -0.5
        X = run(nla, side, uplo, trans, alpha, func, A, B)
        check_bound(side, uplo, trans, alpha, func, A, B, X)
        ref = O.rectrxm(side, uplo, trans, alpha, func, A, B, nthreads=8)
        tol = 50 * n * eps(dt)
        assert O.rel_error(X, ref) < tol, (side, uplo, trans, func, O.rel_error(X, ref))


PANEL_SHAPES = [(256, 16), (512, 32), (192, 32), (1024, 16), (768, 32)]


@pytest.mark.parametrize("rows,strip,waves", [s + (4,) for s in PANEL_SHAPES] + [(512, 16, 8), (512, 32, 8),
                                                                                  (192, 16, 8)])
@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_panel_shapes(nla, rows, strip, waves, dt):
    """Every panel size / strip width / wave count the configuration (and a
    tuner) may pick gives the same results as the oracle: all combos at sizes
    around the panel boundaries."""
    from nextla_amd import _lib
    keys = ((_lib.NLA_CFG_PANEL_ROWS_F64, _lib.NLA_CFG_PANEL_STRIP_F64) if dt == np.float64
            else (_lib.NLA_CFG_PANEL_ROWS_F32, _lib.NLA_CFG_PANEL_STRIP_F32))
    old = (_lib.set_config(keys[0], rows), _lib.set_config(keys[1], strip))
    old_w = _lib.set_config(_lib.NLA_CFG_PANEL_WAVES, waves)
    assert old[0] >= 0 and old[1] > 0 and old_w >= 0  # rows 0: chosen by layout
    try:
        rng = np.random.default_rng(rows + strip)
        for n, m in ((rows, 33), (rows + 70, 50), (3 * rows + 5, 40)):
            for side, uplo, trans, func in itertools.product("LR", "LU", "NT", "SM"):
                A = O._f(np.tril(rng.uniform(-1, 1, (n, n)) / np.sqrt(n)) + np.diag(2 + rng.random(n))).astype(dt)
                if uplo == "U":
                    A = np.asfortranarray(A.T)
                B = rng.standard_normal((n, m) if side == "L" else (m, n)).astype(dt)
                alpha = 1.5 if func == "S" else -0.5
                X = run(nla, side, uplo, trans, alpha, func, A, B)
                check_bound(side, uplo, trans, alpha, func, A, B, X)
    finally:
        _lib.set_config(keys[0], old[0])
        _lib.set_config(keys[1], old[1])
        _lib.set_config(_lib.NLA_CFG_PANEL_WAVES, old_w)


def test_config_rejects_invalid_panel_shapes(nla):
    from nextla_amd import _lib
    assert _lib.set_config(_lib.NLA_CFG_PANEL_ROWS_F64, 100) == -1
    assert _lib.set_config(_lib.NLA_CFG_PANEL_ROWS_F64, 1088) == -1
    assert _lib.set_config(_lib.NLA_CFG_PANEL_STRIP_F32, 24) == -1
    assert _lib.set_config(_lib.NLA_CFG_PANEL_WAVES, 6) == -1
    assert _lib.set_config(99, 1) == -1


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_unit_diagonal(nla, dt):
    """trsm / trmm with diag 'U' (nla_trsm / nla_trmm): the diagonal holds
    NaN and must never be read; results match BLAS with diag 'U' and the
    error bounds of the unit-diagonal matrix, for every combo and sizes that
    take one panel, several panels and the recursion."""
    rng = np.random.default_rng(41)
    for n, m in ((1, 3), (63, 8), (300, 17), (1500, 24)):
        for side, uplo, trans, func in itertools.product("LR", "LU", "NT", "SM"):
            A = np.tril(rng.uniform(-1, 1, (n, n)) / np.sqrt(n))
            if uplo == "U":
                A = A.T.copy()
            A1 = A.copy()
            np.fill_diagonal(A1, 1.0)
            An = A.copy()
            np.fill_diagonal(An, np.nan)
            A1, An = A1.astype(dt), An.astype(dt)
            B = rng.standard_normal((n, m) if side == "L" else (m, n)).astype(dt)
            Bd = dev(B)
            (nla.trsm if func == "S" else nla.trmm)(side, uplo, trans, "U", dev(An), Bd, alpha=0.5)
            X = host(Bd)
            assert np.all(np.isfinite(X)), (n, side, uplo, trans, func)
            check_bound(side, uplo, trans, 0.5, func, A1, B, X)
            ref = O.blas(side, uplo, trans, 0.5, func, An, B, diag="U")
            assert O.rel_error(X, ref) < 50 * n * eps(dt), (n, side, uplo, trans, func, O.rel_error(X, ref))


def test_odd_leading_dimensions(nla):
    """Sub-matrix views with odd lda/ldb take the unaligned (scalar) load path."""
    rng = np.random.default_rng(9)
    for dt in (torch.float64, torch.float32):
        npdt = np.float64 if dt == torch.float64 else np.float32
        for side, uplo, trans, func in itertools.product("LR", "LU", "NT", "SM"):
            n, m = 203, 77
            A0 = np.tril(rng.uniform(-1, 1, (n, n)) / np.sqrt(n)) + np.diag(2 + rng.random(n))
            if uplo == "U":
                A0 = A0.T.copy()
            A0 = A0.astype(npdt)
            rows, cols = (n, m) if side == "L" else (m, n)
            B0 = rng.standard_normal((rows, cols)).astype(npdt)
            lda = n + 2                                  # 205, odd
            ldb = rows + (2 if rows % 2 else 1)          # odd
            Av = torch.empty((n, lda), dtype=dt, device=DEV).t()[:n, :n]
            Bv = torch.empty((cols, ldb), dtype=dt, device=DEV).t()[:rows, :cols]
            Av.copy_(torch.from_numpy(A0))
            Bv.copy_(torch.from_numpy(B0))
            assert Av.stride() == (1, lda) and Bv.stride() == (1, ldb)
            nla.unified_rectrxm_(side, uplo, trans, 1.25, func, Av, Bv)
            check_bound(side, uplo, trans, 1.25, func, A0, B0, host(Bv))


def test_row_major_tensors(nla):
    """Row-major (C-order) torch tensors go through the transpose identities."""
    rng = np.random.default_rng(4)
    n, m = 150, 70
    for side, uplo, trans, func in COMBOS:
        A0 = np.tril(rng.uniform(-1, 1, (n, n)) / np.sqrt(n)) + np.diag(2 + rng.random(n))
        if uplo == "U":
            A0 = A0.T.copy()
        B0 = rng.standard_normal((n, m) if side == "L" else (m, n))
        for a_rm, b_rm in ((True, False), (False, True), (True, True)):
            A = torch.from_numpy(A0.copy()).to(DEV) if a_rm else dev(A0)
            B = torch.from_numpy(B0.copy()).to(DEV) if b_rm else dev(B0)
            nla.unified_rectrxm_(side, uplo, trans, 0.5, func, A, B)
            check_bound(side, uplo, trans, 0.5, func, A0, B0, host(B))


def test_alpha_zero_sets_zero(nla):
    B = dev(np.full((40, 9), np.nan))
    A = dev(np.eye(40))
    nla.unified_rectrxm_("L", "L", "N", 0.0, "S", A, B)
    assert np.all(host(B) == 0.0)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_gemm_entry_points(nla, dt):
    rng = np.random.default_rng(6)
    tol = 10 * eps(dt)
    for M, N, K in ((1, 1, 1), (17, 33, 5), (128, 128, 128), (300, 70, 257), (64, 1000, 64), (5, 7, 0)):
        X = rng.standard_normal((M, K)).astype(dt)
        Y = rng.standard_normal((K, N)).astype(dt)
        C = rng.standard_normal((M, N)).astype(dt)
        Cd = dev(C)
        nla.GEMM_ADD_(dev(X), dev(Y), Cd)
        ref = C.astype(np.float64) + X.astype(np.float64) @ Y.astype(np.float64)
        scale = 1 + np.abs(X).astype(np.float64) @ np.abs(Y).astype(np.float64)
        assert np.all(np.abs(host(Cd) - ref) <= tol * max(K, 1) * scale + tol * np.abs(ref))
        Cd = dev(C)
        nla.GEMM_SUB_(Cd, dev(X), dev(Y))
        ref = C.astype(np.float64) - X.astype(np.float64) @ Y.astype(np.float64)
        assert np.all(np.abs(host(Cd) - ref) <= tol * max(K, 1) * scale + tol * np.abs(ref))
        # transposed operands through the mirror's layout handling
        Cd = dev(C)
        Xt = torch.from_numpy(X.copy()).to(DEV)  # row-major: stored as X^T
        Yt = torch.from_numpy(Y.copy()).to(DEV)
        nla.GEMM_ADD_(Xt, Yt, Cd)
        ref = C.astype(np.float64) + X.astype(np.float64) @ Y.astype(np.float64)
        assert np.all(np.abs(host(Cd) - ref) <= tol * max(K, 1) * scale + tol * np.abs(ref))


def test_non_default_stream(nla):
    rng = np.random.default_rng(8)
    n, m = 700, 200
    A0 = np.tril(rng.uniform(-1, 1, (n, n)) / np.sqrt(n)) + np.diag(2 + rng.random(n))
    B0 = rng.standard_normal((n, m))
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        A, B = dev(A0), dev(B0)
        nla.unified_rectrxm_("L", "L", "N", 1.0, "S", A, B)
    s.synchronize()
    check_bound("L", "L", "N", 1.0, "S", A0, B0, B.cpu().numpy())


def _device_backward_error(side, uplo, trans, alpha, A, B0, X):
    """FP64 backward error on the GPU (torch matmul as the checker)."""
    T = A.double()
    T = torch.tril(T) if uplo == "L" else torch.triu(T)
    if trans != "N":
        T = T.t()
    Xd, Bd = X.double(), B0.double()
    R = alpha * Bd - (T @ Xd if side == "L" else Xd @ T)
    return (torch.linalg.norm(R) / (torch.linalg.norm(T) * torch.linalg.norm(Xd) + abs(alpha) * torch.linalg.norm(Bd))).item()


@pytest.mark.parametrize("dt", [torch.float64, torch.float32])
def test_full_size_8192(nla, dt):
    """n = 8192, B 8192 x 8192, alpha = 1.5 (the drop-in benchmark size):
    backward error <= 10 n eps for left-lower and right-upper-transposed TRSM,
    and the TRMM -> TRSM round trip recovers B."""
    from nextla_amd import inputs
    n = 8192
    npdt = np.float64 if dt == torch.float64 else np.float32
    for side, uplo, trans in (("L", "L", "N"), ("R", "U", "T")):
        A, B = inputs.device_workload1(n, n, uplo, side, dt, seed=17)
        B0 = B.clone()
        nla.unified_rectrxm_(side, uplo, trans, 1.5, "S", A, B)
        torch.cuda.synchronize()
        e = _device_backward_error(side, uplo, trans, 1.5, A, B0, B)
        assert e <= 10 * n * eps(npdt), (side, uplo, trans, e)
        # round trip: TRMM of the solution recovers 1.5*B0
        nla.unified_rectrxm_(side, uplo, trans, 1.0, "M", A, B)
        torch.cuda.synchronize()
        rel = (torch.linalg.norm(B.double() - 1.5 * B0.double()) / torch.linalg.norm(1.5 * B0.double())).item()
        assert rel < 1e3 * eps(npdt), rel
        del A, B, B0
        torch.cuda.empty_cache()
