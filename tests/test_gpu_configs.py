"""The benchmark workloads themselves, at their full sizes, checked against
the FP64 oracle:

  * the plain drop-in at n = 8192 (B 8192 x 8192, alpha = 1.5): every
    side/uplo/trans(N,T) combination x TRSM/TRMM x FP32/FP64.  On the device:
    the normwise backward error (TRSM) or product error (TRMM) in FP64 against
    10*n*eps of the dtype.  On the host: the forward error of 64 sampled
    right-hand sides (columns of B for side 'L', rows for side 'R') against
    OpenBLAS dtrsm / dtrmm in FP64 (the reference tests' oracle,
    test/unified_rectrxm.jl:36-40) on the same A;
  * the FP64-accurate refined solve at n = 16384 with 4096 right-hand sides
    and measured cond_2(A) ~ 1e2, 1e4, 1e6, bf16 and fp16 updates: the
    backward error recomputed on the device in FP64, both normwise
    (Frobenius) and per column (max_j ||r_j|| / (||A||_2 ||x_j|| + ||s_j||),
    with ||A||_2 bounded below by A's largest column norm), and the solution
    against the plain FP64 path and against OpenBLAS on sampled columns;
  * the 10,000-system batch (n log-uniform on [16, 1024], RHS log-uniform on
    [1, 256]) against one plain call per system;
  * diagonal blocks with a large condition number (the leaf applies explicit
    inverses of its 64 x 64 blocks; the bound below must hold anyway).
"""
import math

import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu
DEV = "cuda"
COMBOS8 = [(s, u, t) for s in "LR" for u in "LU" for t in "NT"]


def eps(dt):
    return float(torch.finfo(dt).eps)


def colmajor_empty(rows, cols, dt):
    return torch.empty((cols, rows), dtype=dt, device=DEV).t()


def gen_tri(n, uplo, dt, seed):
    """Off-diagonals U(-1,1)/sqrt(n), diagonal 2+U(0,1), zero outside the
    triangle (the 8192 drop-in workload's A)."""
    g = torch.Generator(device=DEV)
    g.manual_seed(seed)
    a = colmajor_empty(n, n, dt).uniform_(-1.0, 1.0, generator=g).mul_(1.0 / math.sqrt(n))
    a.copy_(torch.tril(a) if uplo == "L" else torch.triu(a))
    a.diagonal().copy_(torch.empty(n, dtype=dt, device=DEV).uniform_(2.0, 3.0, generator=g))
    return a


def gen_rhs(rows, cols, dt, seed):
    g = torch.Generator(device=DEV)
    g.manual_seed(seed)
    return colmajor_empty(rows, cols, dt).normal_(0.0, 1.0, generator=g)


def op_dev(A, uplo, trans):
    T = torch.tril(A.double()) if uplo == "L" else torch.triu(A.double())
    return T.t() if trans != "N" else T


def device_error(side, uplo, trans, alpha, func, A, B0, X):
    """TRSM: ||alpha B0 - op(A) X|| / (||A|| ||X|| + |alpha| ||B0||);
    TRMM: ||X - alpha op(A) B0|| / (|alpha| ||A|| ||B0||); FP64, Frobenius."""
    T = op_dev(A, uplo, trans)
    Xd, Bd = X.double(), B0.double()
    if func == "S":
        R = alpha * Bd - (T @ Xd if side == "L" else Xd @ T)
        den = torch.linalg.norm(T) * torch.linalg.norm(Xd) + abs(alpha) * torch.linalg.norm(Bd)
    else:
        R = Xd - alpha * (T @ Bd if side == "L" else Bd @ T)
        den = abs(alpha) * torch.linalg.norm(T) * torch.linalg.norm(Bd)
    return (torch.linalg.norm(R) / den).item()


_HOST_A = {}


def host_tri(dt, uplo, A):
    """FP64 host copy of a device triangle, cached per (dtype, uplo)."""
    key = (dt, uplo)
    if key not in _HOST_A:
        _HOST_A.clear()
        _HOST_A[key] = np.asfortranarray(A.double().cpu().numpy())
    return _HOST_A[key]


@pytest.mark.parametrize("func", ["S", "M"])
@pytest.mark.parametrize("dt", [torch.float64, torch.float32], ids=["f64", "f32"])
def test_drop_in_8192_all_combos(nla, dt, func):
    n, alpha, ns = 8192, 1.5, 64
    sample = np.random.default_rng(5).choice(n, ns, replace=False)
    # FP64 forward error of the sampled right-hand sides: kappa(A) of this
    # workload is about 2, so the bound leaves room for n-dependent growth
    # of the rounding errors only
    fwd_tol = {torch.float64: 1e-13, torch.float32: 2e-5}[dt]
    worst = {}
    for uplo in "LU":
        A = gen_tri(n, uplo, dt, seed=31 if uplo == "L" else 32)
        Ah = host_tri(dt, uplo, A)
        for side in "LR":
            B0 = gen_rhs(n, n, dt, seed=41)
            for trans in "NT":
                B = B0.clone()
                nla.unified_rectrxm_(side, uplo, trans, alpha, func, A, B)
                torch.cuda.synchronize()
                e = device_error(side, uplo, trans, alpha, func, A, B0, B)
                assert e <= 10 * n * eps(dt), (side, uplo, trans, func, e)
                # sampled right-hand sides against OpenBLAS in FP64
                if side == "L":
                    bs = B0[:, sample].double().cpu().numpy()
                    got = B[:, sample].double().cpu().numpy()
                else:
                    bs = B0[sample, :].double().cpu().numpy()
                    got = B[sample, :].double().cpu().numpy()
                ref = O.blas(side, uplo, trans, alpha, func, Ah, np.asfortranarray(bs))
                fe = O.rel_error(got, ref)
                assert fe <= fwd_tol, (side, uplo, trans, func, fe)
                worst[(side, uplo, trans)] = (e, fe)
                del B
            del B0
        del A
        torch.cuda.empty_cache()
    print(f"\n{dt} {func}: " + ", ".join(f"{''.join(k)} bwd {v[0]:.1e} fwd {v[1]:.1e}" for k, v in worst.items()))


def columnwise_backward_error(side, A, S, X):
    """max_j ||s_j - A x_j|| / (||A||_lb ||x_j|| + ||s_j||), columns of the
    left-side system, with ||A||_lb = max column 2-norm <= ||A||_2 (so the
    value is an upper bound of the error with the exact 2-norm)."""
    T = torch.tril(A)
    R = S - T @ X
    lb = torch.linalg.norm(T, dim=0).max()
    num = torch.linalg.norm(R, dim=0)
    den = lb * torch.linalg.norm(X, dim=0) + torch.linalg.norm(S, dim=0)
    return (num / den).max().item()


_C2 = {}


def config2(cond):
    """A (n = 16384, calibrated to cond_2 ~ cond) and B (4096 RHS), cached."""
    from nextla_amd import inputs
    if cond not in _C2:
        _C2.clear()
        torch.cuda.empty_cache()
        A, B0, c, kappa = inputs.calibrate_precision_tri(16384, 4096, cond, seed=5)
        P = B0.clone()
        import nextla_amd as nla
        nla.unified_rectrxm_("L", "L", "N", 1.0, "S", A, P)
        torch.cuda.synchronize()
        _C2[cond] = (A, B0, P, kappa)
    return _C2[cond]


@pytest.mark.parametrize("prec", ["bf16", "fp16", "fp16x2", "auto"])
@pytest.mark.parametrize("cond", [1e2, 1e4, 1e6])
def test_refined_16384_4096(nla, cond, prec):
    from nextla_amd.mixed import rectrxm_mixed_
    A, B0, P, kappa = config2(cond)
    assert abs(math.log10(kappa) - math.log10(cond)) <= 0.35, kappa
    X = B0.clone()
    _, info = rectrxm_mixed_("L", "L", "N", 1.0, A, X, precision=prec, tol=1e-14)
    assert info["converged"], info
    T = torch.tril(A)
    R = B0 - T @ X
    bwd_f = (torch.linalg.norm(R) / (torch.linalg.norm(T) * torch.linalg.norm(X) + torch.linalg.norm(B0))).item()
    bwd_c = columnwise_backward_error("L", A, B0, X)
    # the library stops on the per-column form; the Frobenius form is smaller
    assert bwd_c <= 1.5e-14, (bwd_c, info)
    assert bwd_f <= bwd_c * 1.01, (bwd_f, bwd_c)
    # the plain FP64 solution of the same system (the refined solution's
    # target) and OpenBLAS on sampled columns: forward errors within
    # kappa * 1e-14
    rel = (torch.linalg.norm(X - P) / torch.linalg.norm(P)).item()
    assert rel <= kappa * 1e-14, (rel, kappa)
    cols = np.random.default_rng(2).choice(4096, 16, replace=False)
    Ah = np.asfortranarray(A.cpu().numpy())
    ref = O.blas("L", "L", "N", 1.0, "S", Ah, np.asfortranarray(B0[:, cols].cpu().numpy()))
    fe = O.rel_error(X[:, cols].cpu().numpy(), ref)
    assert fe <= kappa * 1e-14, (fe, kappa)
    print(f"\ncond {kappa:.2e} {prec}: {info['iterations']} steps ({info['precision']}), bwd col {bwd_c:.2e} "
          f"frob {bwd_f:.2e}, vs plain {rel:.2e}, vs OpenBLAS {fe:.2e}")


def batch_systems(dt, count, seed):
    """The 10,000-system workload: n log-uniform on [16, 1024], RHS
    log-uniform on [1, 256], A lower with off-diagonals U(-1,1)/n and diagonal
    1+U(0,1), B N(0,1)."""
    rng = np.random.default_rng(seed)
    g = torch.Generator(device=DEV)
    g.manual_seed(seed)
    ns = np.exp(rng.uniform(np.log(16), np.log(1024), count)).astype(int)
    ms = np.exp(rng.uniform(np.log(1), np.log(256), count)).astype(int)
    As, Bs = [], []
    for n, m in zip(ns, ms):
        A = colmajor_empty(n, n, dt).uniform_(-1.0, 1.0, generator=g).mul_(1.0 / n)
        A.diagonal().uniform_(1.0, 2.0, generator=g)
        As.append(A)
        Bs.append(colmajor_empty(n, m, dt).normal_(generator=g))
    return As, Bs


@pytest.mark.parametrize("dt", [torch.float64, torch.float32], ids=["f64", "f32"])
def test_batch_10000_systems(nla, dt):
    from nextla_amd.batch import BatchPlan
    As, B0s = batch_systems(dt, 10000, seed=0)
    Xb = [b.clone() for b in B0s]
    Xp = [b.clone() for b in B0s]
    plan = BatchPlan("L", "L", "N", "S", As, Xb, 1.0)
    plan.execute(use_graph=True)
    torch.cuda.synchronize()
    plan.close()
    for A, X in zip(As, Xp):
        nla.unified_rectrxm_("L", "L", "N", 1.0, "S", A, X)
    torch.cuda.synchronize()
    e = eps(dt)
    worst_b, worst_d = 0.0, 0.0
    for A, B0, X, Y in zip(As, B0s, Xb, Xp):
        n = A.shape[0]
        T = torch.tril(A.double())
        Xd = X.double()
        bwd = (torch.linalg.norm(B0.double() - T @ Xd)
               / (torch.linalg.norm(T) * torch.linalg.norm(Xd) + torch.linalg.norm(B0.double()))).item()
        d = (torch.linalg.norm(Xd - Y.double()) / torch.linalg.norm(Y.double())).item()
        # cond(A) of these systems is below 4: the two solutions differ by
        # at most the sum of their forward errors
        assert bwd <= 10 * n * e, (n, X.shape[1], bwd)
        assert d <= 2 * 4 * 10 * n * e, (n, X.shape[1], d)
        worst_b = max(worst_b, bwd / (n * e))
        worst_d = max(worst_d, d / (n * e))
    print(f"\n{dt}: worst backward error {worst_b:.2e} n eps, batch vs plain {worst_d:.2e} n eps")


def ill_block_tri(n, dt, kappa_blk, seed, uplo="L"):
    """Lower (or upper) A whose every 64 x 64 diagonal block has 1-norm
    condition number ~kappa_blk: unit-scale diagonal 1+U(0,1), and inside
    each block off-diagonals U(-1,1)*s with s tuned for the target (strict
    triangles of such matrices make the inverse grow geometrically); the
    blocks below the diagonal are U(-1,1)/n."""
    rng = np.random.default_rng(seed)
    A = np.tril(rng.uniform(-1, 1, (n, n)) / n, -1)
    A[np.diag_indices(n)] = 1.0 + rng.random(n)
    for b0 in range(0, n, 64):
        nb = min(64, n - b0)
        L = np.tril(rng.uniform(-1, 1, (nb, nb)), -1)
        D = np.diag(A[np.arange(b0, b0 + nb), np.arange(b0, b0 + nb)])
        lo, hi = 0.0, 8.0
        for _ in range(50):
            s = 0.5 * (lo + hi)
            k = np.linalg.cond(L * s + D, 1)
            lo, hi = (s, hi) if k < kappa_blk else (lo, s)
        A[b0:b0 + nb, b0:b0 + nb] = L * lo + D
    if uplo == "U":
        A = A.T
    return np.asfortranarray(A.astype(dt))


@pytest.mark.parametrize("dt,kappas", [(np.float64, (1e6, 1e8, 1e10)), (np.float32, (1e3, 1e4, 1e5))],
                         ids=["f64", "f32"])
def test_ill_conditioned_diagonal_blocks(nla, dt, kappas):
    """B = A X_true, so the solution is not dominated by inv(A)'s largest
    direction: the case where applying an explicit inverse loses accuracy
    in proportion to the block's condition number."""
    e = float(np.finfo(dt).eps)
    rng = np.random.default_rng(9)
    for kb in kappas:
        for n in (64, 200, 512, 1100):
            for side, uplo, trans in (("L", "L", "N"), ("L", "U", "T"), ("R", "U", "N"), ("R", "L", "T")):
                A = ill_block_tri(n, dt, kb, seed=int(kb) % 97 + n, uplo=uplo)
                Xt = rng.standard_normal((n, 24) if side == "L" else (24, n))
                Op = O.op_matrix(uplo, trans, A)
                B = (Op @ Xt if side == "L" else Xt @ Op).astype(dt)
                Ad = torch.from_numpy(np.ascontiguousarray(A.T)).to(DEV).t()
                Bd = torch.from_numpy(np.ascontiguousarray(B.T)).to(DEV).t()
                nla.unified_rectrxm_(side, uplo, trans, 1.0, "S", Ad, Bd)
                torch.cuda.synchronize()
                X = Bd.cpu().numpy()
                bwd = O.backward_error(side, uplo, trans, 1.0, A, B, X)
                assert bwd <= 10 * n * e, (kb, n, side, uplo, trans, bwd, 10 * n * e)
