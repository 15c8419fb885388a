"""Recursive LU with partial pivoting on the device (nla_getrf2, the
reference's getrf2!, src/lu.jl:185-300) against the oracle's restatement,
which is itself pinned to LAPACK dgetrf (tests/test_oracle.py).

FP64: the same pivots as the oracle and factors equal to rounding.  FP32:
P*L*U = A to the bound of the reference tests' check (test/lu.jl:25-28,
relative error of the reconstruction) and |L| <= 1; pivots are not compared
in FP32, where a different summation order can flip a near tie.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu
DEV = "cuda"
SHAPES = [(1, 1), (1, 7), (7, 1), (2, 2), (3, 5), (64, 64), (100, 90), (90, 100), (257, 300), (513, 513),
          (1000, 700)]


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a).T)).to(DEV).t()


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("m,n", SHAPES)
def test_fp64_matches_oracle(nla, m, n):
    rng = np.random.default_rng(m * 1009 + n)
    A = rng.standard_normal((m, n))
    LU, ipiv, info = nla.getrf2_(dev(A))
    ref, ref_piv, ref_info = O.getrf2(A)
    assert info == ref_info == 0
    assert np.array_equal(host(ipiv), ref_piv), (m, n)
    got = host(LU)
    assert np.max(np.abs(got - ref)) <= 1e-11 * max(1.0, np.max(np.abs(ref))), (m, n)
    assert np.linalg.norm(O.lu_reconstruct(got, host(ipiv)) - A) <= 20 * max(m, n) * 2.2e-16 * np.linalg.norm(A)


@pytest.mark.parametrize("m,n", [(64, 64), (300, 257), (257, 300), (1000, 1000)])
def test_fp32_reconstructs(nla, m, n):
    rng = np.random.default_rng(7 + m + n)
    A = rng.standard_normal((m, n)).astype(np.float32)
    LU, ipiv, info = nla.getrf2_(dev(A))
    got, piv = host(LU), host(ipiv)
    assert info == 0
    k = min(m, n)
    assert np.all((piv >= np.arange(1, k + 1)) & (piv <= m))
    assert np.max(np.abs(np.tril(got[:, :k], -1))) <= 1.0 + 1e-6
    err = np.linalg.norm(O.lu_reconstruct(got, piv) - A) / np.linalg.norm(A)
    assert err <= 20 * max(m, n) * 1.2e-7, err


def test_singular_info_and_padded_lda(nla):
    """A zero third column: info = 3 (U(3,3) exactly zero) and the rest of
    the factorization as the oracle's; A inside a larger allocation (lda > m)
    with the rows past m untouched."""
    rng = np.random.default_rng(3)
    A = rng.standard_normal((40, 30))
    A[:, 2] = 0.0
    big = torch.full((30, 48), 7.0, dtype=torch.float64, device=DEV).t()  # 48 x 30, lda 48
    big[:40].copy_(torch.from_numpy(A).to(DEV))
    view = big[:40]
    LU, ipiv, info = nla.getrf2_(view)
    ref, ref_piv, ref_info = O.getrf2(A)
    assert info == ref_info == 3
    assert np.array_equal(host(ipiv), ref_piv)
    assert np.max(np.abs(host(LU) - ref)) <= 1e-12 * np.max(np.abs(ref))
    assert torch.all(big[40:] == 7.0)


def test_empty_and_invalid(nla):
    E = torch.empty((0, 5), dtype=torch.float64, device=DEV)
    _, ipiv, info = nla.getrf2_(E)
    assert info == 0 and ipiv.numel() == 0
    with pytest.raises(ValueError):
        nla.getrf2_(torch.ones(4, 5, dtype=torch.float64, device=DEV))  # row-major
    with pytest.raises(ValueError):
        nla.getrf2_(dev(np.eye(4)), ipiv=torch.empty(2, dtype=torch.int64, device=DEV))
    with pytest.raises(ValueError):
        nla.getrf2_(dev(np.eye(4)).to(torch.int32))


def test_large_fp64_self_certified(nla):
    """n = 4096: P*L*U = A checked on the device in FP64."""
    n = 4096
    g = torch.Generator(device=DEV)
    g.manual_seed(5)
    A0 = torch.empty((n, n), dtype=torch.float64, device=DEV).t().normal_(generator=g)
    A = A0.clone()
    LU, ipiv, info = nla.getrf2_(A)
    assert info == 0
    L = torch.tril(LU, -1) + torch.eye(n, dtype=torch.float64, device=DEV)
    U = torch.triu(LU)
    R = L @ U
    perm = torch.arange(n, device=DEV)
    piv = ipiv.cpu().numpy() - 1
    p = perm.cpu().numpy()
    for i in range(n):  # row i of P*A came from row p[i]
        j = piv[i]
        p[i], p[j] = p[j], p[i]
    PA = A0[torch.from_numpy(p).to(DEV)]
    err = (torch.linalg.norm(R - PA) / torch.linalg.norm(A0)).item()
    assert err <= 1e-13, err
    assert torch.max(torch.abs(torch.tril(LU, -1))).item() <= 1.0


@pytest.fixture
def split_rows():
    """Set the row count from which leaves spread over several work-groups;
    restores the previous value."""
    from nextla_amd import _lib
    olds = []

    def set_(v):
        olds.append(_lib.set_config(_lib.NLA_CFG_LU_SPLIT_ROWS, v))
    yield set_
    if olds:
        _lib.set_config(_lib.NLA_CFG_LU_SPLIT_ROWS, olds[0])


@pytest.mark.parametrize("m,n", SHAPES + [(2049, 8), (4097, 20)])
def test_split_leaf_matches_oracle(nla, split_rows, m, n):
    """Every leaf through the multi-work-group kernel (threshold 1 row): the
    same pivots as the oracle and factors equal to rounding, for the shapes
    above and blocks with a partial last work-group."""
    split_rows(1)
    rng = np.random.default_rng(m * 7919 + n)
    A = rng.standard_normal((m, n))
    LU, ipiv, info = nla.getrf2_(dev(A))
    ref, ref_piv, ref_info = O.getrf2(A)
    assert info == ref_info == 0
    assert np.array_equal(host(ipiv), ref_piv), (m, n)
    assert np.max(np.abs(host(LU) - ref)) <= 1e-11 * max(1.0, np.max(np.abs(ref))), (m, n)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("m,n", [(131072, 8), (131073, 8), (6000, 8), (8192, 64)])
def test_split_leaf_equals_single_group_leaf(nla, split_rows, dtype, m, n):
    """Tall leaves up to the 64-work-group limit (131072 rows; one more row
    takes the one-work-group kernel): the split kernel gives the same pivots
    and the same factors as the one-work-group leaf on the same input."""
    g = torch.Generator(device=DEV)
    g.manual_seed(m + n)
    A0 = torch.empty((n, m), dtype=dtype, device=DEV).normal_(generator=g).t()
    split_rows(0)
    L1, p1, i1 = nla.getrf2_(A0.clone())
    split_rows(1)
    L2, p2, i2 = nla.getrf2_(A0.clone())
    assert i1 == i2 == 0
    assert torch.equal(p1, p2)
    tol = 1e-12 if dtype == torch.float64 else 1e-5
    assert torch.max(torch.abs(L1 - L2)).item() <= tol * torch.max(torch.abs(L1)).item()


@pytest.mark.parametrize("rows,split", [(3000, 1), (5000, 1), (5000, 0)],
                         ids=["one-group-register", "split", "global-memory"])
def test_leaf_singular_and_padded(nla, split_rows, rows, split):
    """A zero column and lda > m through each leaf kernel (m = 5000 with
    split rows 1 takes the multi-work-group kernel for its leaves, 3000 the
    one-work-group register kernel, split rows 0 the global-memory leaf):
    info, pivots and factors as the oracle's, the rows past m untouched."""
    split_rows(split)
    rng = np.random.default_rng(11)
    A = rng.standard_normal((rows, 12))
    A[:, 2] = 0.0
    big = torch.full((12, rows + 100), 7.0, dtype=torch.float64, device=DEV).t()
    big[:rows].copy_(torch.from_numpy(A).to(DEV))
    LU, ipiv, info = nla.getrf2_(big[:rows])
    ref, ref_piv, ref_info = O.getrf2(A)
    assert info == ref_info == 3
    assert np.array_equal(host(ipiv), ref_piv)
    assert np.max(np.abs(host(LU) - ref)) <= 1e-12 * np.max(np.abs(ref))
    assert torch.all(big[rows:] == 7.0)


@pytest.mark.parametrize("rows,split", [(5000, 1), (5000, 0), (2000, 1)])
def test_leaf_all_nan_column(nla, split_rows, rows, split):
    """An all-NaN column: no finite pivot, so its row stays (pivot = the row
    itself, as the oracle's first-maximum rule gives for all-NaN), info 0,
    and the NaNs spread exactly where the oracle's do."""
    split_rows(split)
    rng = np.random.default_rng(12)
    A = rng.standard_normal((rows, 10))
    A[:, 4] = np.nan
    LU, ipiv, info = nla.getrf2_(dev(A))
    ref, ref_piv, ref_info = O.getrf2(A)
    got = host(LU)
    assert info == ref_info == 0
    assert np.array_equal(host(ipiv), ref_piv)
    assert np.array_equal(np.isnan(got), np.isnan(ref))
    ok = ~np.isnan(ref)
    assert np.max(np.abs(got[ok] - ref[ok])) <= 1e-12 * np.max(np.abs(ref[ok]))


@pytest.fixture
def spin_limit():
    from nextla_amd import _lib
    olds = []

    def set_(v):
        olds.append(_lib.set_config(_lib.NLA_CFG_LU_SPIN_LIMIT, v))
    yield set_
    if olds:
        _lib.set_config(_lib.NLA_CFG_LU_SPIN_LIMIT, olds[0])


@pytest.mark.parametrize("m,n", [(20000, 8), (9000, 24)])
def test_split_leaf_timeout_is_recovered(nla, split_rows, spin_limit, m, n):
    """A grid-barrier wait of the multi-work-group leaf that runs out (forced
    from the host with a one-poll limit, so work-groups that are merely a
    little behind count as missing): the leaf is redone from its saved copy
    by the one-work-group leaf, and the factorization equals the normal one
    (pivots equal, factors to rounding), singular info included."""
    g = torch.Generator(device=DEV)
    g.manual_seed(m + n)
    A0 = torch.empty((n, m), dtype=torch.float64, device=DEV).normal_(generator=g).t()
    A0[:, 3] = 0.0
    split_rows(1)
    L1, p1, i1 = nla.getrf2_(A0.clone())
    spin_limit(1)
    for _ in range(2):
        L2, p2, i2 = nla.getrf2_(A0.clone())
        assert i1 == i2 == 4
        assert torch.equal(p1, p2)
        assert torch.max(torch.abs(L1 - L2)).item() <= 1e-12 * torch.max(torch.abs(L1)).item()
    spin_limit(0)
    L3, p3, i3 = nla.getrf2_(A0.clone())  # and the normal path after a recovery
    assert i3 == 4 and torch.equal(p1, p3) and torch.equal(L1, L3)
