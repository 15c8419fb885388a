"""Synthetic inputs for the triangular solve / multiply workloads.

Host generators (numpy, seeded PCG64) produce the matrices used by the tests
and fixtures; `device_*` generators build the large benchmark inputs directly
in HBM with torch's seeded device generator.  A and B are returned as
column-major (Fortran-order) arrays, the layout of Julia's Matrix.
"""
import math

import numpy as np


def _rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def tri_mask(a, uplo):
    return np.tril(a) if uplo == "L" else np.triu(a)


def reference_test_tri(n, uplo, rng, dtype=np.float64):
    """`tri(rand(n,n) .+ 1) + 10I` of test/unified_rectrxm.jl:20-25 and
    test/trsm.jl:12: off-diagonals U(1,2), diagonal U(11,12)."""
    a = tri_mask(rng.random((n, n)) + 1.0, uplo) + 10.0 * np.eye(n)
    return np.asfortranarray(a.astype(dtype))


def reference_test_rhs(rows, cols, rng, dtype=np.float64):
    """`rand(rows, cols) .+ 1` of test/unified_rectrxm.jl:27."""
    return np.asfortranarray((rng.random((rows, cols)) + 1.0).astype(dtype))


def scaled_tri(n, uplo, rng, offscale, diag_lo, dtype=np.float64):
    """Off-diagonals U(-1,1)*offscale, diagonal diag_lo + U(0,1)."""
    a = (rng.random((n, n)) * 2.0 - 1.0) * offscale
    a = tri_mask(a, uplo)
    a[np.diag_indices(n)] = diag_lo + rng.random(n)
    return np.asfortranarray(a.astype(dtype))


def workload0_tri(n=512, seed=0):
    """Lower A, off-diagonals U(-1,1)/n, diagonal 1+U(0,1), seed 0 (the
    512 x 512 CPU-runnable case)."""
    return scaled_tri(n, "L", _rng(seed), 1.0 / n, 1.0)


def workload1_tri(n, uplo, rng, dtype=np.float64):
    """Off-diagonals U(-1,1)/sqrt(n), diagonal 2+U(0,1) (the 8192 drop-in case)."""
    return scaled_tri(n, uplo, rng, 1.0 / math.sqrt(n), 2.0, dtype)


def batch_tri(n, rng, dtype=np.float64):
    """Off-diagonals U(-1,1)/n, diagonal 1+U(0,1) (batched systems)."""
    return scaled_tri(n, "L", rng, 1.0 / n, 1.0, dtype)


def normal_rhs(rows, cols, rng, dtype=np.float64):
    return np.asfortranarray(rng.standard_normal((rows, cols)).astype(dtype))


def device_colmajor(rows, cols, dtype, device="cuda"):
    """Uninitialised column-major (rows x cols) device matrix."""
    import torch
    return torch.empty((cols, rows), dtype=dtype, device=device).t()


def device_workload1(n, m, uplo, side, dtype, seed, device="cuda"):
    """A (n x n, off-diagonal U(-1,1)/sqrt(n), diagonal 2+U(0,1), zero outside
    the `uplo` triangle) and B (n x m for side 'L', m x n for 'R', N(0,1)),
    generated in HBM in column-major order."""
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    a = device_colmajor(n, n, dtype, device)
    a.uniform_(-1.0, 1.0, generator=g)
    a.mul_(1.0 / math.sqrt(n))
    if uplo == "L":
        a.copy_(torch.tril(a))
    else:
        a.copy_(torch.triu(a))
    d = torch.empty(n, dtype=dtype, device=device).uniform_(0.0, 1.0, generator=g).add_(2.0)
    a.diagonal().copy_(d)
    rows, cols = (n, m) if side == "L" else (m, n)
    b = device_colmajor(rows, cols, dtype, device)
    b.normal_(0.0, 1.0, generator=g)
    return a, b


def log_diagonal(n, c, dtype=np.float64):
    """Diagonal 2*c^(i/(n-1)), i = 0..n-1: log-spaced from 2 to 2c."""
    t = np.arange(n, dtype=np.float64) / max(n - 1, 1)
    return (2.0 * np.power(float(c), t)).astype(dtype)


def precision_tri(n, c, rng, dtype=np.float64):
    """Lower A for the refinement workload: off-diagonals U(-1,1)/sqrt(n),
    log-spaced diagonal 2 .. 2c (log_diagonal); cond(A) grows about linearly
    with c, and calibrate_precision_tri picks c for a measured target."""
    a = tri_mask((rng.random((n, n)) * 2.0 - 1.0) / math.sqrt(n), "L")
    a[np.diag_indices(n)] = log_diagonal(n, c)
    return np.asfortranarray(a.astype(dtype))


def cond1_estimate(solve, solve_t, anorm1, x0, iters=5):
    """kappa_1(A) = ||A||_1 * ||A^-1||_1 with ||A^-1||_1 from Hager's
    estimator (Hager 1984; Higham 1988, the algorithm of LAPACK's xLACON).
    solve(x) = A^-1 x and solve_t(x) = A^-T x act on float64 column vectors
    (numpy arrays or torch tensors of shape (n, 1)); x0 is the starting
    vector, all entries 1/n.  The estimate is a lower bound, exact for most
    matrices and within a small factor otherwise."""
    x = x0
    est = 0.0
    for k in range(iters):
        y = solve(x)
        est_new = float(abs(y).sum())
        xi = y * 0.0 + 1.0
        xi[y < 0] = -1.0
        z = solve_t(xi)
        zabs = abs(z)
        if float(zabs.max()) <= float((z * x).sum()) or (k > 0 and est_new <= est):
            est = max(est, est_new)
            break
        est = est_new
        j = int(zabs.argmax())
        x = x * 0.0
        x[j, 0] = 1.0
    return anorm1 * est


def cond2_estimate(mv, mv_t, solve, solve_t, x0, iters=40):
    """kappa_2(A) = sigma_max / sigma_min (the 2-norm condition number, the
    meaning of Julia's and numpy's cond(A)) by power iteration: sigma_max^2
    is the largest eigenvalue of A^T A (x <- A^T A x), sigma_min^-2 the
    largest of A^-1 A^-T (x <- A^-1 A^-T x).  mv / mv_t apply A / A^T,
    solve / solve_t apply A^-1 / A^-T to float64 column vectors (numpy arrays
    or torch tensors of shape (n, 1)); x0 is the (nonzero) starting vector.
    Both Rayleigh-quotient estimates grow monotonically towards the true
    values from below."""
    def top(apply):
        x = x0 / float((x0 * x0).sum()) ** 0.5
        lam = 0.0
        for _ in range(iters):
            y = apply(x)
            lam = float((x * y).sum())
            x = y / float((y * y).sum()) ** 0.5
        return lam
    smax2 = top(lambda x: mv_t(mv(x)))
    sinv2 = top(lambda x: solve(solve_t(x)))
    return (smax2 * sinv2) ** 0.5


def device_cond2(A, uplo="L", iters=40, seed=1):
    """kappa_2 of a triangular device matrix (zero outside its triangle):
    products by torch, solves by the library's own TRSM."""
    import torch
    from . import unified_rectrxm_
    n = A.shape[0]

    def solve(x, t="N"):
        y = x.clone()
        unified_rectrxm_("L", uplo, t, 1.0, "S", A, y)
        return y

    g = torch.Generator(device=A.device)
    g.manual_seed(seed)
    x0 = torch.rand((n, 1), dtype=torch.float64, device=A.device, generator=g) + 0.5
    return cond2_estimate(lambda x: A @ x, lambda x: A.t() @ x, solve, lambda x: solve(x, "T"), x0, iters)


def device_cond1(A, uplo="L"):
    """kappa_1 of a triangular device matrix (zero outside its triangle),
    with the solves done by the library's own TRSM."""
    import torch
    from . import unified_rectrxm_
    n = A.shape[0]
    anorm = float(A.abs().sum(0).max())

    def solve(x):
        y = x.clone()
        unified_rectrxm_("L", uplo, "N", 1.0, "S", A, y)
        return y

    def solve_t(x):
        y = x.clone()
        unified_rectrxm_("L", uplo, "T", 1.0, "S", A, y)
        return y

    x0 = torch.full((n, 1), 1.0 / n, dtype=torch.float64, device=A.device)
    return cond1_estimate(solve, solve_t, anorm, x0)


def device_precision_tri(n, m, c, seed, device="cuda"):
    """The refinement workload in HBM: lower A as precision_tri (FP64) with
    diagonal scale c, and B (n x m) N(0,1), column-major."""
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    a = device_colmajor(n, n, torch.float64, device)
    a.uniform_(-1.0, 1.0, generator=g)
    a.mul_(1.0 / math.sqrt(n))
    a.copy_(torch.tril(a))
    a.diagonal().copy_(torch.from_numpy(log_diagonal(n, c)).to(device))
    b = device_colmajor(n, m, torch.float64, device)
    b.normal_(0.0, 1.0, generator=g)
    return a, b


def calibrate_precision_tri(n, m, target, seed, device="cuda", rounds=3, tol=0.15):
    """Choose the diagonal scale c so that the measured kappa_2(A) (the
    2-norm condition number, the meaning of cond(A)) of device_precision_tri
    is within `tol` (relative, in log10 terms: about +-40 % for tol 0.15) of
    `target`.  Returns (A, B, c, measured kappa_2)."""
    c = float(target)
    A = B = None
    kappa = None
    for _ in range(rounds):
        A, B = device_precision_tri(n, m, c, seed, device)
        kappa = device_cond2(A, "L")
        if abs(math.log10(kappa) - math.log10(target)) <= tol:
            break
        c *= target / kappa
    return A, B, c, kappa
