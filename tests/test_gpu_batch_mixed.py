"""Batched FP64-accurate TRSM by fp32 correction solves and FP64 refinement
(nla_batch_mixed_*) against the FP64 oracle.

Each system must reach the requested normwise backward error, recomputed
here in numpy, and agree with OpenBLAS dtrsm (the reference tests' oracle,
test/unified_rectrxm.jl:36-40) within kappa(A) * 1e-14.  A system's result,
error and step count must not depend on the rest of the batch, and hipGraph
replays must match the eager launches bit for bit.
"""
import itertools

import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu
DEV = "cuda"

# (n, m, c): c sets cond(A) through the log-spaced diagonal 2 .. 2c
SHAPES = [(1, 1, 1.0), (17, 3, 10.0), (64, 16, 100.0), (130, 1, 1e3), (300, 40, 5e3), (1024, 9, 1e2),
          (1500, 20, 1e3)]


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a).T)).to(DEV).t()


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def system(rng, n, m, c, side, uplo):
    from nextla_amd import inputs
    A = inputs.precision_tri(n, c, rng)
    if uplo == "U":
        A = np.asfortranarray(A.T)
    B = inputs.normal_rhs(*((n, m) if side == "L" else (m, n)), rng)
    return A, B


@pytest.mark.parametrize("side,uplo,trans", list(itertools.product("LR", "LU", "NT")))
def test_refined_batch_matches_fp64_oracle(nla, side, uplo, trans):
    from nextla_amd import inputs
    from nextla_amd.batch import batched_rectrxm_mixed_
    rng = inputs._rng(31)
    hostA, hostB = [], []
    for n, m, c in SHAPES:
        A, B = system(rng, n, m, c, side, uplo)
        hostA.append(A)
        hostB.append(B)
    alphas = [1.5, -0.5, 2.0, 1.0, 0.75, 1.0, -1.25]
    Bs = [dev(B) for B in hostB]
    _, bwd, its = batched_rectrxm_mixed_(side, uplo, trans, [dev(A) for A in hostA], Bs, alphas=alphas, tol=1e-14)
    for A, B, X, al, b, k in zip(hostA, hostB, Bs, alphas, bwd, its):
        Xh = host(X)
        e = O.backward_error(side, uplo, trans, al, A, B, Xh)
        # the device's stopping test, and an independent numpy evaluation
        # (the two differ by the rounding of the product at this level)
        assert b <= 1e-14 and e <= 1.5e-14, (A.shape[0], b, e, k)
        ref = O.blas(side, uplo, trans, al, "S", A, B)
        kappa = np.linalg.cond(O.op_matrix(uplo, trans, A), 1)
        assert O.rel_error(Xh, ref) <= kappa * 1e-14, (A.shape[0], O.rel_error(Xh, ref), kappa)
        assert 0 <= k <= 4, k


def test_alone_equals_in_mix_and_graph(nla):
    """Bitwise: X, error and step count of a system alone, inside a mixed
    batch, and from hipGraph replays of that batch after refilling B."""
    from nextla_amd import inputs
    from nextla_amd.batch import MixedBatchPlan
    rng = inputs._rng(9)
    A, B = system(rng, 700, 33, 2e3, "L", "L")
    Xa = dev(B)
    pa = MixedBatchPlan("L", "L", "N", [dev(A)], [Xa], 1.25)
    ba, ia, _ = pa.execute()
    others = [system(rng, n, m, c, "L", "L") for n, m, c in [(40, 500, 10.0), (2000, 5, 5e3), (90, 2, 1.0)]]
    As = [dev(a) for a, _ in others] + [dev(A)]
    Bh = [b for _, b in others] + [B]
    Bs = [dev(b) for b in Bh]
    pm = MixedBatchPlan("L", "L", "N", As, Bs, [1.0, -1.0, 3.0, 1.25])
    bm, im, _ = pm.execute()
    assert torch.equal(Xa, Bs[-1]) and ba[0] == bm[-1] and ia[0] == im[-1], (ba, bm, ia, im)
    eager = [b.clone() for b in Bs]
    for rep in range(3):
        for t, b in zip(Bs, Bh):
            t.copy_(dev(b))
        bg, ig, _ = pm.execute(use_graph=True)
        torch.cuda.synchronize()
        assert bg == bm and ig == im, rep
        for t, e in zip(Bs, eager):
            assert torch.equal(t, e), rep


def test_alpha_zero_empty_and_not_converged(nla):
    from nextla_amd import inputs
    from nextla_amd.batch import MixedBatchPlan
    rng = inputs._rng(4)
    A, B = system(rng, 200, 6, 10.0, "L", "L")
    A2, B2 = system(rng, 50, 3, 10.0, "L", "L")
    X0 = dev(B)
    empty_a = torch.empty((0, 0), dtype=torch.float64, device=DEV)
    empty_b = torch.empty((0, 5), dtype=torch.float64, device=DEV)
    plan = MixedBatchPlan("L", "L", "N", [dev(A), empty_a, dev(A2)], [X0, empty_b, dev(B2)], [0.0, 1.0, 1.0])
    bwd, its, conv = plan.execute()
    assert all(conv) and bwd[0] == 0.0 and bwd[1] == 0.0 and its[1] == 0, (bwd, its)
    assert np.all(host(X0) == 0.0)
    # no correction allowed: the fp32 solve's error is reported (and is far
    # above rounding level, so the device and numpy evaluations agree)
    X = dev(B2)
    plan = MixedBatchPlan("L", "L", "N", [dev(A2)], [X])
    bwd, its, conv = plan.execute(tol=1e-14, max_iter=0, raise_on_stall=False)
    e = O.backward_error("L", "L", "N", 1.0, A2, B2, host(X))
    assert not conv[0] and its[0] == 0 and 1e-14 < e < 1e-5, (bwd, e)
    assert abs(e - bwd[0]) <= 1e-3 * e, (bwd, e)
    with pytest.raises(nla.NLAError):
        plan.execute(tol=1e-14, max_iter=0)


def test_rejects_fp32(nla):
    from nextla_amd.batch import MixedBatchPlan
    A = torch.eye(8, device=DEV)
    with pytest.raises(ValueError):
        MixedBatchPlan("L", "L", "N", [A], [torch.ones(8, 2, device=DEV).t().contiguous().t()])
