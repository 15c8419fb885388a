"""Mixed-precision TRSM with FP64 iterative refinement (nla_rectrxm_mixed)
against the FP64 oracle.

Every case must reach the requested normwise backward error (computed
independently here in FP64 with numpy) and agree with OpenBLAS dtrsm (the
reference tests' oracle, test/unified_rectrxm.jl:36-40) within
kappa(A) * 1e-14.  The low-precision part is pinned separately: with no
refinement step the bf16 / fp16 solves must be measurably less accurate
than fp32 (so the low-precision updates really ran) and still close (so
they computed the right product).
"""
import itertools

import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu
DEV = "cuda"
COMBOS = list(itertools.product("LR", "LU", "NT"))
PREC_NAMES = ("bf16", "fp16", "fp32", "fp16x2")


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a).T)).to(DEV).t()


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def case(n, m, side, uplo, c, seed):
    from nextla_amd import inputs
    rng = inputs._rng(seed)
    A = inputs.precision_tri(n, c, rng)
    if uplo == "U":
        A = np.asfortranarray(A.T)
    rows, cols = (n, m) if side == "L" else (m, n)
    B = inputs.normal_rhs(rows, cols, rng)
    return A, B


@pytest.mark.parametrize("side,uplo,trans", COMBOS)
@pytest.mark.parametrize("prec", ["bf16", "fp16", "fp32", "fp16x2", "auto"])
def test_refined_matches_fp64_oracle(nla, side, uplo, trans, prec):
    from nextla_amd.mixed import rectrxm_mixed_
    n, m, alpha = 1100, 37, 1.5  # n > 512: the recursion makes low-precision updates
    A, B = case(n, m, side, uplo, 100.0, seed=7)
    Bd = dev(B)
    _, info = rectrxm_mixed_(side, uplo, trans, alpha, dev(A), Bd, precision=prec, tol=1e-14)
    X = host(Bd)
    bwd = O.backward_error(side, uplo, trans, alpha, A, B, X)
    assert info["converged"] and info["backward_error"] <= 1e-14, info
    # the library's stopping test and an independent numpy evaluation agree
    assert bwd <= 1.5e-14, (bwd, info)
    ref = O.blas(side, uplo, trans, alpha, "S", A, B)
    kappa = np.linalg.cond(O.op_matrix(uplo, trans, A), 1)
    assert O.rel_error(X, ref) <= kappa * 1e-14, (O.rel_error(X, ref), kappa)
    # a well-conditioned system converges in the requested precision (auto:
    # in one of them, within as few steps as fp32 needs)
    assert info["precision"] == prec or (prec == "auto" and info["precision"] in PREC_NAMES), info
    assert info["iterations"] <= {"bf16": 10, "fp16": 6, "fp32": 3, "fp16x2": 3, "auto": 3}[prec], info
    h = info["history"]
    assert all(h[i + 1] < h[i] for i in range(len(h) - 1)), h


def test_low_precision_solve_accuracy(nla):
    """max_iter = 0 returns the first correction solve.  Its backward error
    follows the update precision (bf16 u = 2^-8, fp16 u = 2^-11, fp16 pairs
    about 2^-22, fp32 u = 2^-24): each step down costs accuracy, and none is
    far off (which would mean a wrong product rather than a rounded one).
    The fp16 pairs land near fp32, far from fp16: a pair product that lost
    its hi*lo or lo*hi term would fall back to fp16's error."""
    from nextla_amd.mixed import rectrxm_mixed_
    n, m = 2048, 64
    A, B = case(n, m, "L", "L", 100.0, seed=11)
    bwd = {}
    for prec in ("bf16", "fp16", "fp16x2", "fp32"):
        Bd = dev(B)
        _, info = rectrxm_mixed_("L", "L", "N", 1.0, dev(A), Bd, precision=prec, tol=1e-14, max_iter=0,
                                 raise_on_stall=False)
        X = host(Bd)
        bwd[prec] = O.backward_error("L", "L", "N", 1.0, A, B, X)
        assert info["iterations"] == 0
        # the whole-matrix form, computed here in numpy, and the library's
        assert abs(bwd[prec] - info["backward_error_frobenius"]) <= 1e-3 * bwd[prec] + 1e-18, (bwd, info)
        assert info["backward_error"] >= info["backward_error_frobenius"], info
    assert bwd["bf16"] < 1e-3, bwd
    assert 3.0 < bwd["bf16"] / bwd["fp16"] < 40.0, bwd   # 3 bits apart
    assert bwd["fp16"] / bwd["fp32"] > 100.0, bwd          # 13 bits apart
    assert bwd["fp16"] / bwd["fp16x2"] > 100.0, bwd        # about 11 bits apart
    assert bwd["fp16x2"] / bwd["fp32"] < 16.0, bwd         # about 2 bits apart


@pytest.mark.parametrize("side,uplo,trans", COMBOS)
def test_large_update_tiles_every_orientation(nla, side, uplo, trans):
    """bf16 / fp16 updates of at least 1024 x 1024 run in 256 x 256 tiles,
    here with ragged edges (n = 2300, 1100 right-hand sides) and every operand
    orientation.  With no refinement
    step each solve keeps its precision's accuracy relative to fp32, so every
    tile computed the right product.  Here the fp32 panels set the floor
    (fp16 lands 60-100 times above fp32), and the fp16 pairs sit on it."""
    from nextla_amd.mixed import rectrxm_mixed_
    n, m = 2300, 1100
    A, B = case(n, m, side, uplo, 100.0, seed=19)
    bwd = {}
    for prec in ("bf16", "fp16", "fp16x2", "fp32"):
        Bd = dev(B)
        _, info = rectrxm_mixed_(side, uplo, trans, 1.0, dev(A), Bd, precision=prec, tol=1e-14, max_iter=0,
                                 raise_on_stall=False)
        bwd[prec] = O.backward_error(side, uplo, trans, 1.0, A, B, host(Bd))
    assert bwd["bf16"] < 1e-3, bwd
    assert 3.0 < bwd["bf16"] / bwd["fp16"] < 40.0, bwd
    assert bwd["fp16"] / bwd["fp16x2"] > 20.0, bwd
    assert bwd["fp16x2"] / bwd["fp32"] < 4.0, bwd


def test_not_converged_reports_last_iterate(nla):
    from nextla_amd.mixed import rectrxm_mixed_
    A, B = case(700, 5, "L", "L", 100.0, seed=3)
    Bd = dev(B)
    _, info = rectrxm_mixed_("L", "L", "N", 1.0, dev(A), Bd, precision="bf16", tol=1e-40, max_iter=2,
                             raise_on_stall=False)
    assert not info["converged"] and info["iterations"] == 2 and len(info["history"]) == 3
    X = host(Bd)
    bwd = O.backward_error("L", "L", "N", 1.0, A, B, X)
    assert abs(bwd - info["backward_error_frobenius"]) <= 1e-2 * bwd + 1e-18
    with pytest.raises(nla.NLAError):
        rectrxm_mixed_("L", "L", "N", 1.0, dev(A), dev(B), precision="bf16", tol=1e-40, max_iter=1)


def test_alpha_zero_and_empty(nla):
    from nextla_amd.mixed import rectrxm_mixed_
    A, B = case(600, 4, "L", "L", 10.0, seed=5)
    Bd = dev(B)
    _, info = rectrxm_mixed_("L", "L", "N", 0.0, dev(A), Bd, precision="fp16")
    assert info["converged"] and np.all(host(Bd) == 0.0)
    E = torch.empty((600, 0), dtype=torch.float64, device=DEV)
    _, info = rectrxm_mixed_("L", "L", "N", 1.0, dev(A), E, precision="fp16")
    assert info["iterations"] == 0


def test_rejects_invalid_arguments(nla):
    from nextla_amd.mixed import rectrxm_mixed_
    A, B = case(64, 4, "L", "L", 10.0, seed=5)
    with pytest.raises(ValueError):
        rectrxm_mixed_("L", "L", "N", 1.0, dev(A).float(), dev(B).float())
    with pytest.raises(ValueError):
        rectrxm_mixed_("L", "L", "N", 1.0, dev(A), dev(B), precision="int8")
    with pytest.raises(ValueError):
        rectrxm_mixed_("L", "L", "N", 1.0, dev(A), dev(B), tol=0.0)


@pytest.mark.parametrize("prec", ["bf16", "fp16", "auto"])
def test_full_size_self_certified(nla, prec):
    """n = 16384 at cond_1 ~ 1e4 (fewer right-hand sides than the benchmark):
    the backward error is recomputed here on the device in FP64 and the
    solution compared with the plain FP64 path."""
    from nextla_amd import inputs
    from nextla_amd.mixed import rectrxm_mixed_
    n, m = 16384, 256
    A, B0 = inputs.device_precision_tri(n, m, 1e4 / 2.0, seed=21)
    B = B0.clone()
    _, info = rectrxm_mixed_("L", "L", "N", 1.0, A, B, precision=prec, tol=1e-14)
    assert info["converged"], info
    R = B0 - torch.tril(A) @ B
    bwd = (torch.linalg.norm(R) / (torch.linalg.norm(torch.tril(A)) * torch.linalg.norm(B)
                                   + torch.linalg.norm(B0))).item()
    assert bwd <= 1.5e-14, (bwd, info)
    P = B0.clone()
    nla.unified_rectrxm_("L", "L", "N", 1.0, "S", A, P)
    rel = (torch.linalg.norm(B - P) / torch.linalg.norm(P)).item()
    assert rel <= 1e-10, rel


@pytest.mark.parametrize("side,uplo,trans", [("L", "L", "N"), ("R", "U", "T"), ("L", "U", "T"), ("R", "L", "N")])
@pytest.mark.parametrize("residual", ["int8", "fp64"])
def test_fp64_route_certifies_the_plain_solve(nla, side, uplo, trans, residual):
    """precision 'fp64': the plain FP64 solve certified by one residual.  With
    the FP64 residual X is the plain solve bit for bit; with the int8 one it
    is that solve rounded to >= 55 bits per system.  No correction is needed
    (iterations 0, precision 'fp64')."""
    from nextla_amd import _lib
    from nextla_amd.mixed import rectrxm_mixed_
    n, m, alpha = 2304, 96, -0.75  # n >= 2048, m >= 64: the int8 residual applies
    A, B = case(n, m, side, uplo, 1e3, seed=13)
    Ad = dev(A)
    P = dev(B)
    nla.unified_rectrxm_(side, uplo, trans, alpha, "S", Ad, P)
    old = _lib.set_config(_lib.NLA_CFG_RESIDUAL, _lib.NLA_RES_INT8 if residual == "int8" else _lib.NLA_RES_FP64)
    try:
        Bd = dev(B)
        _, info = rectrxm_mixed_(side, uplo, trans, alpha, Ad, Bd, precision="fp64", tol=1e-14)
    finally:
        _lib.set_config(_lib.NLA_CFG_RESIDUAL, old)
    assert info["converged"] and info["iterations"] == 0 and info["precision"] == "fp64", info
    assert info["residual"] == residual
    X, Xp = host(Bd), host(P)
    if residual == "fp64":
        assert np.array_equal(X, Xp)
    else:
        sysmax = np.abs(Xp).max(0 if side == "L" else 1, keepdims=True)
        assert np.all(np.abs(X - Xp) <= 2.0 ** -55 * sysmax)
    assert O.backward_error(side, uplo, trans, alpha, A, B, X) <= 1e-14
