"""LAUUM on the device (nla_lauum, the reference's lauum!, src/lauum.jl:52)
against the oracle (pinned to LAPACK dlauum): both triangles, FP32 / FP64,
sizes around the 64-row leaves and the recursion, the other triangle and the
rows past n untouched, and the reference test's acceptance bound
(test/lauum.jl: norm(A - expected) / n < rtol)."""
import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu
DEV = "cuda"


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a).T)).to(DEV).t()


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("uplo", ["U", "L"])
@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_matches_oracle(nla, uplo, dt):
    rng = np.random.default_rng(3 if uplo == "U" else 4)
    eps = np.finfo(dt).eps
    for n in (1, 16, 63, 64, 65, 127, 128, 129, 192, 200, 257, 513, 1000):  # around the 64- and 128-row leaves
        A = rng.uniform(-0.5, 0.5, (n + 3, n)).astype(dt)  # 3 extra rows: lda > n
        X = dev(A)
        nla.lauum_(uplo, n, X)
        got = host(X)
        ref = O.lauum(uplo, n, A)
        scale = max(1.0, np.max(np.abs(ref)))
        assert np.max(np.abs(got - ref)) <= 50 * n * eps * scale, (uplo, n)
        assert np.array_equal(got[n:], A[n:])
        assert np.linalg.norm(got - ref) / max(n, 1) < (1e-5 if dt == np.float32 else 1e-12)


def test_rejects_invalid(nla):
    A = dev(np.eye(4))
    with pytest.raises(ValueError):
        nla.lauum_("X", 4, A)
    with pytest.raises(ValueError):
        nla.lauum_("U", -1, A)
    with pytest.raises(ValueError):
        nla.lauum_("U", 5, A)
    with pytest.raises(ValueError):
        nla.lauum_("U", 4, torch.ones(4, 4, dtype=torch.float64, device=DEV))  # row-major


def test_large_self_certified(nla):
    """n = 4096 FP64 lower: the triangle of L^T L recomputed on the device."""
    n = 4096
    g = torch.Generator(device=DEV)
    g.manual_seed(8)
    A0 = torch.empty((n, n), dtype=torch.float64, device=DEV).t().uniform_(-0.5, 0.5, generator=g)
    A = A0.clone()
    nla.lauum_("L", n, A)
    L = torch.tril(A0)
    ref = torch.tril(L.t() @ L)
    got = torch.tril(A)
    assert (torch.linalg.norm(got - ref) / torch.linalg.norm(ref)).item() <= 1e-14
    assert torch.equal(torch.triu(A, 1), torch.triu(A0, 1))
