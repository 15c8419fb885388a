"""Block reflector application on the device (nla_larfb, the reference's
larfb!, src/larfb.jl:37-142) against the oracle, whose H = I - V T V^T is
pinned to LAPACK's QR factor (tests/test_oracle.py).  V and T come from a
Householder QR as in the reference test (test/larfb.jl:14-27); V keeps R in
its upper part and R's diagonal on its diagonal, which must not be read."""
import itertools

import numpy as np
import pytest
import scipy.linalg
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu
DEV = "cuda"


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a).T)).to(DEV).t()


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def reflector(rng, rows, k):
    qr, tau, _, _ = scipy.linalg.lapack.dgeqrf(rng.standard_normal((rows, k)))
    V_unit = np.tril(qr, -1) + np.eye(rows, k)
    return qr, V_unit, O.larft(V_unit, tau)


@pytest.mark.parametrize("side,trans", list(itertools.product("LR", "NTC")))
@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_matches_oracle(nla, side, trans, dt):
    rng = np.random.default_rng(5)
    tol = 1e-12 if dt == np.float64 else 2e-5
    for m, n, k in ((5, 4, 4), (10, 8, 4), (20, 15, 4), (300, 200, 32), (700, 513, 64), (129, 1000, 65)):
        rows = m if side == "L" else n
        k = min(k, rows)
        qr, V_unit, T = reflector(rng, rows, k)
        C = rng.standard_normal((m, n)).astype(dt)
        Cd = dev(C)
        nla.larfb_(side, trans, "F", "C", dev(qr.astype(dt)), dev(T.astype(dt)), Cd)
        ref = O.larfb(side, trans, V_unit.astype(dt), T.astype(dt), C)
        got = host(Cd)
        assert np.max(np.abs(got - ref)) <= tol * max(1.0, np.max(np.abs(ref))) * np.sqrt(max(m, n)), (m, n, k)


def test_padded_and_invalid(nla):
    rng = np.random.default_rng(6)
    qr, V_unit, T = reflector(rng, 40, 8)
    C = rng.standard_normal((40, 30))
    big = torch.full((30, 50), 3.0, dtype=torch.float64, device=DEV).t()  # 50 x 30, ldc 50
    big[:40].copy_(torch.from_numpy(C).to(DEV))
    work = torch.zeros((8, 64), dtype=torch.float64, device=DEV).t()     # ldwork 64 > n = 30
    nla.larfb_("L", "T", "F", "C", dev(qr), dev(T), big[:40], work=work)
    assert np.allclose(host(big[:40]), O.larfb("L", "T", V_unit, T, C), atol=1e-12)
    assert torch.all(big[40:] == 3.0)
    with pytest.raises(ValueError):
        nla.larfb_("L", "N", "B", "C", dev(qr), dev(T), dev(C))
    with pytest.raises(ValueError):
        nla.larfb_("L", "N", "F", "R", dev(qr), dev(T), dev(C))
    with pytest.raises(ValueError):
        nla.larfb_("L", "N", "F", "C", dev(qr).float(), dev(T), dev(C))
