"""The FP64-accurate triangular residual R = alpha*B - op(A)*X (nla_residual),
int8 modular products against the exact-arithmetic oracle:

  * every side/uplo/trans at an order that pads (n = 300): X rounded bit for
    bit as the oracle rounds it, R within a few units of FP64 rounding of the
    exact residual of the rounded operands;
  * X close to the solution (R ~1e-16 of |A||X|), where the rounding of A
    decides R's leading digits: against the same exact model, and against
    the exact residual of the unrounded A within the 2^-bA bound;
  * several row panels and column tiles (n up to 4100): against the FP64
    MFMA product within FP64's own rounding bound;
  * inputs outside the moduli's range or not finite: the FP64 product.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from oracle import residual_oracle as RO

pytestmark = pytest.mark.gpu
DEV = "cuda"
COMBOS8 = [(s, u, t) for s in "LR" for u in "LU" for t in "NT"]


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=np.float64).T)).to(DEV).t()


def empty_like_cm(a):
    return torch.empty((a.shape[1], a.shape[0]), dtype=torch.float64, device=DEV).t()


def tri(n, uplo, rng):
    A = np.tril(rng.uniform(-1, 1, (n, n)) / np.sqrt(n)) + np.diag(2 + rng.random(n))
    return A if uplo == "L" else A.T.copy()


def run(side, uplo, trans, alpha, A, X, B, method):
    from nextla_amd.mixed import residual
    Ad, Xd, Bd = dev(A), dev(X), dev(B)
    Rd = empty_like_cm(Bd)
    residual(side, uplo, trans, alpha, Ad, Xd, Bd, Rd, method=method)
    torch.cuda.synchronize()
    return Rd.cpu().numpy(), Xd.cpu().numpy()


def scale_bound(side, uplo, trans, A, X, bA, bX):
    """2^-100 of rowmax(Ahat) * colmax(Xhat): the reconstruction's absolute
    error floor, in R's shape."""
    Op = O.op_matrix(uplo, trans, A)
    Ah = Op if side == "L" else Op.T
    Xh = X if side == "L" else X.T
    b = 2.0 ** -100 * np.abs(Ah).max(1)[:, None] * np.abs(Xh).max(0)[None, :]
    return b if side == "L" else b.T


@pytest.mark.parametrize("side,uplo,trans", COMBOS8)
def test_int8_residual_matches_exact_model(nla, side, uplo, trans):
    rng = np.random.default_rng(3)
    n, m = 300, 21
    A = tri(n, uplo, rng)
    shp = (n, m) if side == "L" else (m, n)
    X, B = rng.standard_normal(shp), rng.standard_normal(shp)
    X[:, 0 if side == "L" else 0] *= 1e-3  # systems of different scales
    R, Xr = run(side, uplo, trans, 1.5, A, X, B, "int8")
    T, bA, bX = RO.plan_bits(n)
    Rm, Xm = RO.fixed_point_model(side, uplo, trans, 1.5, A, X, B, bA, bX)
    assert np.array_equal(Xr, Xm)
    err = np.abs(R - Rm)
    assert np.all(err <= 4 * 2.0 ** -53 * np.abs(Rm) + scale_bound(side, uplo, trans, A, X, bA, bX)), err.max()


@pytest.mark.parametrize("side,uplo,trans,n", [("L", "L", "N", 12), ("R", "U", "T", 12), ("L", "U", "N", 300),
                                                ("R", "L", "T", 300)])
def test_int8_reconstruction_rounds_once(nla, side, uplo, trans, n):
    """The reconstruction forms the integer residual exactly (base-2^40 limbs
    in exact double arithmetic) and rounds it once: every entry lies within
    3 ulp of the exact fixed-point residual, with no absolute floor.  n = 12
    and 300 take 15 and 16 moduli (the other residual tests reach 17)."""
    rng = np.random.default_rng(n)
    m = 7
    A = tri(n, uplo, rng)
    shp = (n, m) if side == "L" else (m, n)
    X, B = rng.standard_normal(shp), rng.standard_normal(shp)
    R, Xr = run(side, uplo, trans, -0.75, A, X, B, "int8")
    T, bA, bX = RO.plan_bits(n)
    assert T == (15 if n == 12 else 16)
    Rm, Xm = RO.fixed_point_model(side, uplo, trans, -0.75, A, X, B, bA, bX)
    assert np.array_equal(Xr, Xm)
    assert np.all(np.abs(R - Rm) <= 3 * 2.0 ** -53 * np.abs(Rm)), np.abs(R - Rm).max()


@pytest.mark.parametrize("side,uplo,trans", [("L", "L", "N"), ("L", "U", "T"), ("R", "L", "N"), ("R", "U", "T")])
def test_int8_residual_near_solution(nla, side, uplo, trans):
    """X = the plain FP64 solve: R is ~1e-16 of |A||X|, so it is made of the
    solve's rounding errors and of A's and X's fixed-point rounding."""
    rng = np.random.default_rng(5)
    n, m = 260, 9
    A = tri(n, uplo, rng)
    shp = (n, m) if side == "L" else (m, n)
    B = rng.standard_normal(shp)
    Xd = dev(B)
    nla.unified_rectrxm_(side, uplo, trans, 1.0, "S", dev(A), Xd)
    X = Xd.cpu().numpy()
    R, Xr = run(side, uplo, trans, 1.0, A, X, B, "int8")
    T, bA, bX = RO.plan_bits(n)
    Rm, Xm = RO.fixed_point_model(side, uplo, trans, 1.0, A, X, B, bA, bX)
    assert np.array_equal(Xr, Xm)
    assert np.all(np.abs(R - Rm) <= 4 * 2.0 ** -53 * np.abs(Rm) + scale_bound(side, uplo, trans, A, X, bA, bX))
    # the exact residual of the returned X with the unrounded A differs by
    # A's rounding only: 2^-bA of each row's largest entry times |x|
    ex = RO.exact_residual(side, uplo, trans, 1.0, A, Xr, B)
    Op = O.op_matrix(uplo, trans, A)
    Ah = Op if side == "L" else Op.T
    Xh = Xr if side == "L" else Xr.T
    bound = 2.0 ** -bA * np.abs(Ah).max(1)[:, None] * np.abs(Xh).sum(0)[None, :]
    d = np.abs(R - ex)
    assert np.all((d if side == "L" else d.T) <= bound + 4 * 2.0 ** -53 * np.abs(ex if side == "L" else ex.T))
    # and R itself is that small: the solve's backward error
    assert np.abs(R).max() <= 1e-14 * np.abs(B).max()


def gpu_tri(n, uplo, seed):
    g = torch.Generator(device=DEV)
    g.manual_seed(seed)
    a = torch.empty((n, n), dtype=torch.float64, device=DEV).t()
    a.uniform_(-1.0, 1.0, generator=g).mul_(n ** -0.5)
    a.copy_(torch.tril(a) if uplo == "L" else torch.triu(a))
    a.diagonal().add_(2.0)
    return a


@pytest.mark.parametrize("side,uplo,trans,n,m", [("L", "L", "N", 2500, 300), ("L", "U", "T", 4100, 130),
                                                 ("R", "U", "N", 2049, 257), ("R", "L", "T", 3000, 64)])
def test_int8_residual_against_fp64_large(nla, side, uplo, trans, n, m):
    from nextla_amd.mixed import residual
    A = gpu_tri(n, uplo, 7)
    shp = (n, m) if side == "L" else (m, n)
    g = torch.Generator(device=DEV)
    g.manual_seed(8)
    B = torch.empty((shp[1], shp[0]), dtype=torch.float64, device=DEV).t().normal_(generator=g)
    X = B.clone()
    nla.unified_rectrxm_(side, uplo, trans, 1.0, "S", A, X)
    X.add_(torch.empty_like(X).normal_(generator=g), alpha=1e-9)  # an iterate near the solution
    X8 = X.clone()
    R8 = torch.empty_like(B)
    residual(side, uplo, trans, 1.25, A, X8, B, R8, method="int8")
    R64 = torch.empty_like(B)
    residual(side, uplo, trans, 1.25, A, X8, B, R64, method="fp64")  # FP64 product of the same (rounded) X
    torch.cuda.synchronize()
    assert torch.all(torch.isfinite(R8))
    Op = (torch.tril(A) if uplo == "L" else torch.triu(A))
    Op = Op.t() if trans != "N" else Op
    mag = (Op.abs() @ X8.abs()) if side == "L" else (X8.abs() @ Op.abs())
    mag = mag + 1.25 * B.abs()
    # the FP64 product's error is below n u |A||X| elementwise; the int8 one
    # is far smaller, so the two agree to within the FP64 bound
    assert torch.all((R8 - R64).abs() <= (n + 2) * 2.0 ** -53 * mag)
    # X was rounded to >= 55 bits of each system's largest entry
    sysmax = X.abs().amax(0 if side == "L" else 1, keepdim=True)
    assert torch.all((X8 - X).abs() <= 2.0 ** -55 * sysmax)


def test_int8_residual_falls_back_outside_range(nla):
    """alpha*B far above |A||X| (beyond the moduli's range) and a NaN in X:
    R is the FP64 product's."""
    rng = np.random.default_rng(9)
    n, m = 200, 6
    A = tri(n, "L", rng)
    X, B = rng.standard_normal((n, m)), rng.standard_normal((n, m))
    Bbig = B * 1e60
    R8, _ = run("L", "L", "N", 1.0, A, X, Bbig, "int8")
    R64, _ = run("L", "L", "N", 1.0, A, X, Bbig, "fp64")
    assert np.array_equal(R8, R64)
    # a NaN in system 2: that system is left as it was, the others rounded
    # as the model rounds them, and R is the FP64 product's for that X
    Xn = X.copy()
    Xn[5, 2] = np.nan
    R8, X8 = run("L", "L", "N", 1.0, A, Xn, B, "int8")
    T, bA, bX = RO.plan_bits(n)
    _, Xm = RO.fixed_point_model("L", "L", "N", 1.0, A, np.nan_to_num(Xn), B, bA, bX)
    assert np.array_equal(X8[:, 2], Xn[:, 2], equal_nan=True)
    others = [j for j in range(m) if j != 2]
    assert np.array_equal(X8[:, others], Xm[:, others])
    R64, _ = run("L", "L", "N", 1.0, A, X8, B, "fp64")
    assert np.array_equal(R8, R64, equal_nan=True)
    assert np.all(np.isnan(R8[:, 2]))


def test_int8_residual_alpha_zero_and_zero_system(nla):
    rng = np.random.default_rng(10)
    n, m = 300, 5
    A = tri(n, "U", rng)
    X, B = rng.standard_normal((m, n)), rng.standard_normal((m, n))
    X[3, :] = 0.0
    R, Xr = run("R", "U", "N", 0.0, A, X, B, "int8")
    T, bA, bX = RO.plan_bits(n)
    Rm, Xm = RO.fixed_point_model("R", "U", "N", 0.0, A, X, B, bA, bX)
    assert np.array_equal(Xr, Xm)
    assert np.all(np.abs(R - Rm) <= 4 * 2.0 ** -53 * np.abs(Rm) + scale_bound("R", "U", "N", A, X, bA, bX))
    assert np.all(R[3, :] == 0.0)


@pytest.mark.parametrize("side,uplo,trans", [("L", "L", "N"), ("R", "U", "T")])
def test_refined_solve_int8_and_fp64_residuals(nla, side, uplo, trans):
    """The refined solve with either residual converges to the FP64 solution;
    the int8 residual is used where the shape allows it."""
    from nextla_amd import _lib
    from nextla_amd.mixed import rectrxm_mixed_
    n, m = 2304, 96
    A = gpu_tri(n, uplo, 11)
    shp = (n, m) if side == "L" else (m, n)
    g = torch.Generator(device=DEV)
    g.manual_seed(12)
    B = torch.empty((shp[1], shp[0]), dtype=torch.float64, device=DEV).t().normal_(generator=g)
    P = B.clone()
    nla.unified_rectrxm_(side, uplo, trans, 1.0, "S", A, P)
    out = {}
    for method in (_lib.NLA_RES_INT8, _lib.NLA_RES_FP64):
        old = _lib.set_config(_lib.NLA_CFG_RESIDUAL, method)
        try:
            X = B.clone()
            _, info = rectrxm_mixed_(side, uplo, trans, 1.0, A, X, precision="bf16", tol=1e-14)
        finally:
            _lib.set_config(_lib.NLA_CFG_RESIDUAL, old)
        assert info["converged"] and info["backward_error"] <= 1e-14, info
        out[info["residual"]] = X
        rel = (torch.linalg.norm(X - P) / torch.linalg.norm(P)).item()
        assert rel <= 1e-13, (info["residual"], rel)
    assert set(out) == {"int8", "fp64"}


@pytest.mark.parametrize("variant", [1, 2, 3])
@pytest.mark.parametrize("side,uplo,trans,n,m", [("L", "L", "N", 2304, 600), ("R", "U", "T", 2500, 300),
                                                 ("L", "U", "N", 3000, 2100)])
def test_product_schedules_identical(nla, variant, side, uplo, trans, n, m):
    """Every schedule of the modular products (prefetch depth, one counter in
    row order or per-XCD counters over super-tiles, partial super-tiles at
    the edges) gives the same residual and the same rounded X, bit for bit."""
    from nextla_amd import _lib
    from nextla_amd.mixed import residual
    A = gpu_tri(n, uplo, 21)
    shp = (n, m) if side == "L" else (m, n)
    g = torch.Generator(device=DEV)
    g.manual_seed(22)
    B = torch.empty((shp[1], shp[0]), dtype=torch.float64, device=DEV).t().normal_(generator=g)
    X = B.clone().mul_(0.5).add_(torch.empty_like(B).normal_(generator=g), alpha=1e-6)
    out = []
    for v in (0, variant):
        old = _lib.set_config(_lib.NLA_CFG_OZ_VARIANT, v)
        try:
            Xv, Rv = X.clone(), torch.empty_like(B)
            residual(side, uplo, trans, 1.0, A, Xv, B, Rv, method="int8")
            torch.cuda.synchronize()
        finally:
            _lib.set_config(_lib.NLA_CFG_OZ_VARIANT, old)
        out.append((Xv, Rv))
    assert torch.equal(out[0][0], out[1][0])
    assert torch.equal(out[0][1], out[1][1])
