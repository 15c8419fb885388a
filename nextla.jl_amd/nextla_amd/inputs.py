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
