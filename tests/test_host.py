"""Host logic of the Python mirror, on the CPU: the layout handling (column-
major, row-major, padded leading dimensions) and the argument mapping onto
the C ABI.  The device library is replaced by a stand-in that executes the
oracle on the very memory the mirror hands to the C call, so a wrong side /
uplo / trans / n / m / ld mapping shows up as a wrong answer."""
import ctypes
import itertools

import numpy as np
import pytest
import torch

from oracle import oracle as O


def _wrap(ptr, ld, rows, cols, dtype):
    ct = ctypes.c_double if dtype == 1 else ctypes.c_float
    isz = ctypes.sizeof(ct)
    count = ld * (cols - 1) + rows if rows and cols else 0
    buf = (ct * max(count, 1)).from_address(ptr)
    flat = np.ctypeslib.as_array(buf)
    return np.lib.stride_tricks.as_strided(flat, shape=(rows, cols), strides=(isz, isz * ld))


class FakeLib:
    """Executes nla_rectrxm / nla_gemm with the oracle, in place."""

    def nla_rectrxm(self, side, uplo, trans, func, dtype, alpha, n, m, A, lda, B, ldb, stream):
        side, uplo, trans, func = (c.decode() for c in (side, uplo, trans, func))
        Av = _wrap(A, lda, n, n, dtype)
        rows, cols = (n, m) if side == "L" else (m, n)
        Bv = _wrap(B, ldb, rows, cols, dtype)
        Bv[...] = O.rectrxm(side, uplo, trans, alpha, func, Av, Bv)
        return 0

    def nla_gemm(self, dtype, tx, ty, M, N, K, alpha, X, ldx, Y, ldy, beta, C, ldc, stream):
        tx, ty = tx.decode(), ty.decode()
        Xv = _wrap(X, ldx, *((K, M) if tx == "T" else (M, K)), dtype)
        Yv = _wrap(Y, ldy, *((N, K) if ty == "T" else (K, N)), dtype)
        Cv = _wrap(C, ldc, M, N, dtype)
        opx = Xv.T if tx == "T" else Xv
        opy = Yv.T if ty == "T" else Yv
        Cv[...] = beta * Cv + alpha * (opx.astype(np.float64) @ opy.astype(np.float64)).astype(Cv.dtype)
        return 0

    def nla_last_error(self):
        return b"fake"


@pytest.fixture
def mirror(monkeypatch):
    import nextla_amd
    from nextla_amd import _lib
    monkeypatch.setattr(_lib, "load", lambda path=None: FakeLib())
    monkeypatch.setattr(nextla_amd, "_require_device", lambda *ts: None)
    monkeypatch.setattr(nextla_amd, "_stream", lambda t: None)
    return nextla_amd


def _tensor(a, layout, dt):
    a = np.array(a, dtype=dt, copy=True)  # never share memory with the caller's array
    r, c = a.shape
    if layout == "col":
        return torch.from_numpy(np.ascontiguousarray(a.T)).t()
    if layout == "row":
        return torch.from_numpy(np.ascontiguousarray(a))
    # padded column-major view: ld = r + 3
    big = torch.zeros((c, r + 3), dtype=torch.float64 if dt == np.float64 else torch.float32).t()
    v = big[:r, :c]
    v.copy_(torch.from_numpy(a))
    return v


@pytest.mark.parametrize("la,lb", list(itertools.product(("col", "row", "pad"), repeat=2)))
def test_layouts_map_to_correct_call(mirror, la, lb):
    rng = np.random.default_rng(1)
    n, m = 37, 11
    for side, uplo, trans, func in itertools.product("LR", "LU", "NTC", "SM"):
        A = np.tril(rng.random((n, n)) + 1) + 10 * np.eye(n)
        A = A if uplo == "L" else A.T.copy()
        B = rng.standard_normal((n, m) if side == "L" else (m, n))
        At, Bt = _tensor(A, la, np.float64), _tensor(B, lb, np.float64)
        out = mirror.unified_rectrxm_(side, uplo, trans, 1.5, func, At, Bt)
        assert out is Bt
        ref = O.blas(side, uplo, trans, 1.5, func, A, B)
        assert O.rel_error(Bt.numpy(), ref) < 1e-13, (side, uplo, trans, func, la, lb)


@pytest.mark.parametrize("lx,ly,lc", list(itertools.product(("col", "row"), repeat=3)))
def test_gemm_layouts(mirror, lx, ly, lc):
    rng = np.random.default_rng(2)
    X, Y, C = rng.standard_normal((9, 5)), rng.standard_normal((5, 7)), rng.standard_normal((9, 7))
    Ct = _tensor(C, lc, np.float64)
    mirror.GEMM_ADD_(_tensor(X, lx, np.float64), _tensor(Y, ly, np.float64), Ct)
    np.testing.assert_allclose(Ct.numpy(), C + X @ Y, rtol=1e-13)
    Ct = _tensor(C, lc, np.float64)
    mirror.GEMM_SUB_(Ct, _tensor(X, lx, np.float64), _tensor(Y, ly, np.float64))
    np.testing.assert_allclose(Ct.numpy(), C - X @ Y, rtol=1e-13)


def test_shape_errors(mirror):
    A = torch.zeros(4, 4, dtype=torch.float64)
    with pytest.raises(ValueError):
        mirror.unified_rectrxm_("L", "L", "N", 1.0, "S", A, torch.zeros(5, 2, dtype=torch.float64))
    with pytest.raises(ValueError):
        mirror.unified_rectrxm_("R", "L", "N", 1.0, "S", A, torch.zeros(4, 2, dtype=torch.float64))
    with pytest.raises(ValueError):
        mirror.unified_rectrxm_("L", "L", "N", 1.0, "S", torch.zeros(4, 3, dtype=torch.float64),
                                torch.zeros(4, 2, dtype=torch.float64))
    with pytest.raises(ValueError):  # no unit stride
        mirror.unified_rectrxm_("L", "L", "N", 1.0, "S", A, torch.zeros(8, 6, dtype=torch.float64)[::2, ::2])
    with pytest.raises(ValueError):
        mirror.GEMM_ADD_(torch.zeros(3, 4, dtype=torch.float64), torch.zeros(5, 2, dtype=torch.float64),
                         torch.zeros(3, 2, dtype=torch.float64))


def test_panel_rows_by_layout_setting(nla):
    """NLA_CFG_PANEL_ROWS_*: 0 selects the panel size by layout (the
    default), a multiple of 64 in [64, 1024] fixes it, anything else is
    rejected and leaves the setting unchanged."""
    from nextla_amd import _lib
    for key in (_lib.NLA_CFG_PANEL_ROWS_F32, _lib.NLA_CFG_PANEL_ROWS_F64):
        before = _lib.set_config(key, 0)
        assert before == 0 or (before % 64 == 0 and 64 <= before <= 1024)
        try:
            assert _lib.set_config(key, 32) == -1
            assert _lib.set_config(key, 256) == 0
            assert _lib.set_config(key, 0) == 256
        finally:
            _lib.set_config(key, before)


def test_tuner_picks_fastest_and_restores(nla):
    """The tuner's selection and configuration handling, with stand-in timings
    (no GPU): it returns the fastest candidate and leaves the process-wide
    configuration as it found it until apply() is called."""
    from nextla_amd import _lib, tune
    before = _lib.set_config(_lib.NLA_CFG_PANEL_ROWS_F64, 512)
    _lib.set_config(_lib.NLA_CFG_PANEL_ROWS_F64, before)
    fake = {(256, 16): 3.0, (256, 32): 4.0, (384, 16): 2.5, (384, 32): 2.9, (512, 16): 2.6, (512, 32): 5.0}
    res = tune.tune_panels("f64", measure=lambda r, s: fake[(r, s)])
    assert (res["rows"], res["strip"]) == (384, 16)
    assert res["ms"]["384/16"] == 2.5
    assert _lib.set_config(_lib.NLA_CFG_PANEL_ROWS_F64, before) == before
    bres = tune.tune_batch("f64", measure=lambda mk: {0x8: 5.0, 0xC: 4.0, 0xE: 3.0, 0xF: 3.5, 0x1E: 3.2,
                                                      0x1F: 3.1}[mk])
    assert bres["mask"] == 0xE and len(bres["ms"]) == len(tune.BATCH_MASKS)
    old = tune.apply({"panel": {"f64": {"rows": 384, "strip": 16}}, "batch_classes": 0xE})
    try:
        assert _lib.set_config(_lib.NLA_CFG_PANEL_ROWS_F64, 384) == 384
        assert _lib.set_config(_lib.NLA_CFG_BATCH_CLASSES, 0xE) == 0xE
    finally:
        tune.apply({"panel": old["panel"], "batch_classes": old["batch_classes"]})
    assert _lib.set_config(_lib.NLA_CFG_PANEL_ROWS_F64, before) == before
