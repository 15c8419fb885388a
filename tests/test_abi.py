"""The C-ABI library loads, exports every symbol include/nla.h declares, and
validates arguments before any device work (no GPU needed: these calls return
before the first HIP call)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "nla.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(nla_\w+)\s*\(", txt)))


def test_header_lists_entry_points():
    fns = header_functions()
    for f in ("nla_rectrxm", "nla_gemm_add", "nla_gemm_sub", "nla_gemm", "nla_leaf"):
        assert f in fns


def test_library_exports_every_declared_symbol(nla):
    from nextla_amd import _lib
    lib = _lib.load()
    for name in header_functions():
        assert hasattr(lib, name), name
    assert set(header_functions()) == set(_lib.SIGNATURES)
    assert lib.nla_abi_version() >= 1


def _rect(lib, *args):
    side, uplo, trans, func, dtype, alpha, n, m, A, lda, B, ldb = args
    return lib.nla_rectrxm(side, uplo, trans, func, dtype, alpha, n, m, A, lda, B, ldb, None)


GOOD = [b"L", b"L", b"N", b"S", 1, 1.0, 8, 4, 16, 8, 32, 8]


@pytest.mark.parametrize("pos,val", [(0, b"X"), (1, b"Q"), (2, b"Z"), (3, b"K"), (4, 7), (6, -1), (7, -2),
                                     (8, None), (9, 3), (10, None), (11, 2)])
def test_rectrxm_rejects_invalid_arguments(nla, pos, val):
    from nextla_amd import _lib
    lib = _lib.load()
    args = list(GOOD)
    args[pos] = val
    st = _rect(lib, *args)
    assert st == -(pos + 1), (pos, st)
    assert b"invalid" in lib.nla_last_error()


def test_rectrxm_empty_is_noop(nla):
    from nextla_amd import _lib
    lib = _lib.load()
    for n, m in ((0, 4), (8, 0), (0, 0)):
        assert _rect(lib, b"R", b"U", b"T", b"M", 0, 2.0, n, m, None, 1, None, 1) == 0


def test_right_side_ldb_rule(nla):
    from nextla_amd import _lib
    lib = _lib.load()
    # side 'R': B is m x n, so ldb must be >= m (here m=10 > n=8)
    st = _rect(lib, b"R", b"L", b"N", b"S", 1, 1.0, 8, 10, 16, 8, 32, 9)
    assert st == -12


def test_gemm_and_leaf_validation(nla):
    from nextla_amd import _lib
    lib = _lib.load()
    assert lib.nla_gemm(3, b"N", b"N", 4, 4, 4, 1.0, 8, 4, 8, 4, 1.0, 8, 4, None) == -1
    assert lib.nla_gemm(1, b"X", b"N", 4, 4, 4, 1.0, 8, 4, 8, 4, 1.0, 8, 4, None) == -2
    assert lib.nla_gemm(1, b"N", b"N", 4, 4, 4, 1.0, 8, 3, 8, 4, 1.0, 8, 4, None) == -9
    assert lib.nla_gemm(1, b"N", b"T", 4, 5, 4, 1.0, 8, 4, 8, 4, 1.0, 8, 4, None) == -11
    assert lib.nla_gemm(1, b"N", b"N", 0, 4, 4, 1.0, None, 1, None, 1, 1.0, None, 1, None) == 0
    assert lib.nla_gemm_add(1, 4, 4, 4, 8, 3, 8, 4, 8, 4, None) == -6
    assert lib.nla_gemm_sub(1, 4, 4, 4, 8, 2, 8, 4, 8, 4, None) == -6
    assert lib.nla_leaf(b"Q", b"L", b"L", 1, 4, 4, 8, 4, 8, 4, None) == -1
    assert lib.nla_leaf(b"S", b"L", b"L", 1, 4, 4, 8, 4, 8, 3, None) == -10


def test_python_mirror_validates_before_device_work(nla):
    import torch
    A = torch.zeros(4, 4, dtype=torch.float64)
    B = torch.zeros(4, 2, dtype=torch.float64)
    with pytest.raises(ValueError):
        nla.unified_rectrxm_("X", "L", "N", 1.0, "S", A, B)
    with pytest.raises(ValueError):
        nla.unified_rectrxm_("L", "L", "N", 1.0, "S", A, B.to(torch.float32))
    with pytest.raises(ValueError):  # host tensors are rejected, never computed on the CPU
        nla.unified_rectrxm_("L", "L", "N", 1.0, "S", A, B)
    with pytest.raises(ValueError):
        nla.trsm("L", "L", "N", "X", A, B)


def test_batch_validation(nla):
    import ctypes
    from nextla_amd import _lib
    lib = _lib.load()
    items = (_lib.BatchItem * 2)()
    items[0].n, items[0].m, items[0].A, items[0].lda, items[0].B, items[0].ldb, items[0].alpha = 8, 4, 64, 8, 64, 8, 1.0
    items[1].n, items[1].m, items[1].A, items[1].lda, items[1].B, items[1].ldb, items[1].alpha = 70000, 4, 64, 70000, 64, 70000, 1.0
    need = lib.nla_batch_workspace_bytes(b"S", 1, items, 1)
    assert need > 0
    plan = ctypes.c_void_p()
    assert lib.nla_batch_plan_create(b"X", b"L", b"N", b"S", 1, items, 1, 256, need, None, ctypes.byref(plan)) == -1
    assert lib.nla_batch_plan_create(b"L", b"L", b"N", b"S", 1, items, 2, 256, need, None, ctypes.byref(plan)) == -6
    assert lib.nla_batch_plan_create(b"L", b"L", b"N", b"S", 1, items, 1, 256, need - 1, None,
                                     ctypes.byref(plan)) == -9
    assert lib.nla_batch_plan_create(b"L", b"L", b"N", b"S", 1, items, 1, 264, need, None, ctypes.byref(plan)) == -8
    assert lib.nla_batch_plan_create(b"L", b"L", b"N", b"S", 1, items, 1, 256, need, None, None) == -11
    assert lib.nla_batch_plan_destroy(None) == 0


def test_mixed_validation(nla):
    import ctypes
    from nextla_amd import _lib
    lib = _lib.load()
    need = lib.nla_mixed_workspace_bytes(b"L", 100, 8)
    assert need >= 4 * 100 * 100 + 20 * 100 * 8
    assert lib.nla_mixed_workspace_bytes(b"Q", 100, 8) == -1
    info = _lib.MixedInfo()
    good = [b"L", b"L", b"N", 1, 1.0, 100, 8, 256, 100, 512, 100, 1e-14, 10, 1024, need, None, ctypes.byref(info)]

    def call(pos, val):
        args = list(good)
        args[pos] = val
        return lib.nla_rectrxm_mixed(*args)

    assert call(0, b"X") == -1
    assert call(3, 7) == -4
    assert call(3, 6) == -4
    assert call(3, -1) == -4
    assert call(4, float("nan")) == -5
    assert call(5, -1) == -6
    assert call(11, 0.0) == -12
    assert call(12, 31) == -13
    assert call(8, 99) == -9
    assert call(10, 50) == -11
    assert call(13, 1032) == -14
    assert call(14, need - 1) == -15
    assert call(6, 0) == 0 and info.iterations == 0  # empty: nothing to do


def test_trsm_trmm_diag_validation(nla):
    from nextla_amd import _lib
    lib = _lib.load()
    for fn in (lib.nla_trsm, lib.nla_trmm):
        assert fn(b"L", b"L", b"N", b"Q", 1, 1.0, 8, 4, 16, 8, 32, 8, None) == -4
        assert fn(b"L", b"L", b"Z", b"U", 1, 1.0, 8, 4, 16, 8, 32, 8, None) == -3
        assert fn(b"R", b"U", b"T", b"U", 1, 1.0, 8, 10, 16, 8, 32, 9, None) == -12
        assert fn(b"L", b"L", b"N", b"U", 1, 1.0, 0, 4, None, 1, None, 1, None) == 0


def test_batch_mixed_validation(nla):
    import ctypes
    from nextla_amd import _lib
    lib = _lib.load()
    items = (_lib.BatchItem * 2)()
    items[0].n, items[0].m, items[0].A, items[0].lda, items[0].B, items[0].ldb, items[0].alpha = 100, 8, 256, 100, 512, 100, 1.0
    items[1].n, items[1].m, items[1].A, items[1].lda, items[1].B, items[1].ldb, items[1].alpha = 70000, 4, 64, 70000, 64, 70000, 1.0
    need = lib.nla_batch_mixed_workspace_bytes(b"L", items, 1)
    assert need >= 4 * 100 * 100 + 20 * 100 * 8
    assert lib.nla_batch_mixed_workspace_bytes(b"Q", items, 1) == -1
    assert lib.nla_batch_mixed_workspace_bytes(b"L", items, 2) == -1
    plan = ctypes.c_void_p()
    good = [b"L", b"L", b"N", items, 1, 256, need, None, ctypes.byref(plan)]

    def call(pos, val):
        args = list(good)
        args[pos] = val
        return lib.nla_batch_mixed_plan_create(*args)

    assert call(0, b"X") == -1
    assert call(1, b"Q") == -2
    assert call(2, b"Z") == -3
    assert call(4, 2) == -4          # n = 70000
    assert call(5, 264) == -6        # alignment
    assert call(6, need - 1) == -7
    assert call(8, None) == -9
    items[0].ldb = 99
    assert call(4, 1) == -4          # ldb < n
    bwd = (ctypes.c_double * 1)()
    its = (ctypes.c_int32 * 1)()
    assert lib.nla_batch_mixed_plan_execute(None, 1e-14, 10, None, 0, bwd, its) == -1
    assert lib.nla_batch_mixed_plan_destroy(None) == 0


def test_lu_split_rows_config(nla):
    """NLA_CFG_LU_SPLIT_ROWS: register leaves for every LU leaf by default
    (1), 0 selects the global-memory leaf, negative values are rejected; the
    setter returns the previous value.  Unknown keys return -1."""
    from nextla_amd import _lib
    old = _lib.set_config(_lib.NLA_CFG_LU_SPLIT_ROWS, 4096)
    try:
        assert old == 1
        assert _lib.set_config(_lib.NLA_CFG_LU_SPLIT_ROWS, -1) == -1
        assert _lib.set_config(_lib.NLA_CFG_LU_SPLIT_ROWS, 0) == 4096
        assert _lib.set_config(_lib.NLA_CFG_LU_SPLIT_ROWS, 1) == 0
    finally:
        _lib.set_config(_lib.NLA_CFG_LU_SPLIT_ROWS, old)
    assert _lib.set_config(99, 1) == -1
    # the split leaf's spin limit: 0 (default) restores the built-in limit
    assert _lib.set_config(_lib.NLA_CFG_LU_SPIN_LIMIT, -5) == -1
    assert _lib.set_config(_lib.NLA_CFG_LU_SPIN_LIMIT, 7) == 0
    assert _lib.set_config(_lib.NLA_CFG_LU_SPIN_LIMIT, 0) == 7


def test_lu_lauum_larfb_validation(nla):
    """Argument checks of the widened entry points return -k before any
    device work."""
    import ctypes
    from nextla_amd import _lib
    lib = _lib.load()
    p = 4096  # never dereferenced: every call below fails validation first
    assert lib.nla_getrf2(3, 4, 4, p, 4, p, p, None) == -1
    assert lib.nla_getrf2(1, -1, 4, p, 4, p, p, None) == -2
    assert lib.nla_getrf2(1, 4, -1, p, 4, p, p, None) == -3
    assert lib.nla_getrf2(1, 4, 4, None, 4, p, p, None) == -4
    assert lib.nla_getrf2(1, 4, 4, p, 3, p, p, None) == -5
    assert lib.nla_getrf2(1, 4, 4, p, 4, None, p, None) == -6
    assert lib.nla_getrf2(1, 4, 4, p, 4, p, None, None) == -7
    need = lib.nla_lauum_workspace_bytes(1, 200)
    assert need == 128 * 128 * 8 and lib.nla_lauum_workspace_bytes(1, 64) == 8
    assert lib.nla_lauum_workspace_bytes(7, 10) == -1
    assert lib.nla_lauum(b"X", 1, 4, p, 4, None, 0, None) == -1
    assert lib.nla_lauum(b"L", 9, 4, p, 4, None, 0, None) == -2
    assert lib.nla_lauum(b"L", 1, -1, p, 4, None, 0, None) == -3
    assert lib.nla_lauum(b"L", 1, 4, None, 4, None, 0, None) == -4
    assert lib.nla_lauum(b"L", 1, 4, p, 3, None, 0, None) == -5
    assert lib.nla_lauum(b"L", 1, 200, p, 200, p, need - 1, None) == -7
    assert lib.nla_lauum(b"U", 1, 0, None, 1, None, 0, None) == 0
    good = [b"L", b"N", b"F", b"C", 1, 10, 8, 4, p, 10, p, 4, p, 10, p, 8, None]

    def larfb(pos, val):
        args = list(good)
        args[pos] = val
        return lib.nla_larfb(*args)

    assert larfb(0, b"Q") == -1
    assert larfb(1, b"Q") == -2
    assert larfb(2, b"B") == -3      # backward: not supported, as the reference
    assert larfb(3, b"R") == -4      # rowwise: not supported, as the reference
    assert larfb(4, 5) == -5
    assert larfb(7, 11) == -8        # more reflectors than rows of V
    assert larfb(9, 9) == -10        # ldv < m
    assert larfb(11, 3) == -12       # ldt < k
    assert larfb(13, 9) == -14       # ldc < m
    assert larfb(15, 7) == -16       # ldwork < n
    assert larfb(7, 0) == 0          # k = 0: nothing to apply


def test_residual_validation(nla):
    """nla_residual: argument checks in order, workspace sizes, method and
    order limits (no device work)."""
    from nextla_amd import _lib
    lib = _lib.load()
    assert lib.nla_residual_workspace_bytes(b"L", _lib.NLA_RES_FP64, 100, 8) == 256
    w8 = lib.nla_residual_workspace_bytes(b"L", _lib.NLA_RES_INT8, 300, 21)
    # 16 moduli at n = 300 (padded to 512): A planes 3 * 65536 each, X and
    # product planes 256 * 512 each
    assert w8 >= 16 * (3 * 65536 + 2 * 256 * 512)
    assert lib.nla_residual_workspace_bytes(b"R", _lib.NLA_RES_INT8, 65537, 8) == -1
    assert lib.nla_residual_workspace_bytes(b"Q", _lib.NLA_RES_INT8, 10, 8) == -1
    assert lib.nla_residual_workspace_bytes(b"L", 7, 10, 8) == -1
    good = [b"L", b"L", b"N", _lib.NLA_RES_INT8, 1.0, 300, 21, 4096, 300, 8192, 300, 12288, 300, 16384, 300,
            1 << 20, w8, None]

    def call(pos, val):
        args = list(good)
        args[pos] = val
        return lib.nla_residual(*args)

    for pos, val in ((0, b"X"), (1, b"Q"), (2, b"Z"), (3, 5), (4, float("inf")), (5, -1), (6, -1), (7, None),
                     (8, 299), (9, None), (10, 299), (11, None), (12, 299), (13, None), (14, 299), (15, 1000),
                     (16, w8 - 1)):
        assert call(pos, val) == -(pos + 1), (pos, val)
    assert call(5, 0) == 0 and call(6, 0) == 0  # empty: nothing to do
    # the right side checks leading dimensions against m rows
    args = list(good)
    args[0] = b"R"
    args[8] = args[10] = args[12] = args[14] = 21
    args[16] = lib.nla_residual_workspace_bytes(b"R", _lib.NLA_RES_INT8, 300, 21)
    args[5], args[6] = 300, 21
    args[8] = 20
    assert lib.nla_residual(*args) == -9
    assert _lib.set_config(_lib.NLA_CFG_RESIDUAL, 3) == -1
    old = _lib.set_config(_lib.NLA_CFG_RESIDUAL, _lib.NLA_RES_FP64)
    assert old == _lib.NLA_RES_INT8
    small = lib.nla_mixed_workspace_bytes(b"L", 4096, 128)
    _lib.set_config(_lib.NLA_CFG_RESIDUAL, old)
    assert lib.nla_mixed_workspace_bytes(b"L", 4096, 128) > small + 16 * 4096 * 4096 // 2


def test_serve_create_validation(nla):
    """nla_serve_create rejects bad arguments before touching the device."""
    from nextla_amd import _lib
    lib = _lib.load()
    srv = ctypes.c_void_p()
    assert lib.nla_serve_create(0, 1, 1, 0, 0, 0, ctypes.byref(srv)) == -1
    assert lib.nla_serve_create(1 << 20, 0, 1, 0, 0, 0, ctypes.byref(srv)) == -2
    assert lib.nla_serve_create(1 << 20, 1, 0, 0, 0, 0, ctypes.byref(srv)) == -3
    assert lib.nla_serve_create(1 << 20, 1, 1, 2, 0, 0, ctypes.byref(srv)) == -4
    assert lib.nla_serve_create(1 << 20, 1, 1, _lib.NLA_SERVE_REFINED, 0, 0, ctypes.byref(srv)) == -5
    assert lib.nla_serve_create(1 << 20, 1, 1, 0, 0, 0, None) == -7
    assert lib.nla_serve_destroy(None) == 0
    assert lib.nla_serve_step(None, None, 0, None) == -1
