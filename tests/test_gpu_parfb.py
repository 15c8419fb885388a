"""pamm! / parfb! (tile QR's triangular-pentagonal block reflectors) on the
device, through the C ABI, against the oracle's restatement of
src/pamm.jl / src/parfb.jl and the xTPRFB closed form it is pinned to:

  * parfb for every side/trans/direct/storev (the reference's test grid,
    test/parfb.jl:5-46, which skips left-forward-rowwise; it is included
    here), FP32 and FP64, the reference's shape and larger ones, l = 0 and
    l = k;
  * pamm 'W' and 'A' on their own for every side/storev/direct;
  * quick returns, argument errors, padded leading dimensions, repeatability
    (test/pamm.jl:85-162).
"""
import numpy as np
import pytest
import torch

from oracle import tprfb_oracle as TO

pytestmark = pytest.mark.gpu
DEV = "cuda"
COMBOS = [(s, t, d, v) for s in "LR" for t in "NT" for d in "FB" for v in "CR"]


def dev(a, dtype, pad=0):
    """Column-major device copy of a, with `pad` extra rows of leading
    dimension (filled with a sentinel)."""
    a = np.asarray(a)
    r, c = a.shape
    buf = torch.full((c, r + pad), 7.25, dtype=dtype, device=DEV)
    buf[:, :r] = torch.from_numpy(np.ascontiguousarray(a.T)).to(dtype)
    return buf.t()[:r]


def host(t):
    return t.double().cpu().numpy()


def rel(a, b):
    d = np.linalg.norm(a - b)
    s = max(np.linalg.norm(b), 1e-300)
    return d / s


TOL = {torch.float64: 1e-12, torch.float32: 2e-5}


@pytest.mark.parametrize("dt", [torch.float64, torch.float32])
@pytest.mark.parametrize("side,trans,direct,storev", COMBOS)
@pytest.mark.parametrize("m2,n2,k,l", [(20, 24, 8, 4), (9, 7, 5, 0), (6, 9, 6, 6), (300, 257, 64, 32),
                                       (700, 513, 64, 64)])
def test_parfb_matches_oracle(nla, dt, side, trans, direct, storev, m2, n2, k, l):
    from nextla_amd import parfb_
    rng = np.random.default_rng(m2 * 7 + n2 * 3 + k + l)
    if l > (m2 if side == "L" else n2):
        pytest.skip("triangle larger than V")
    m1, n1 = (k, n2) if side == "L" else (m2, k)
    A1 = rng.standard_normal((m1, n1))
    A2 = rng.standard_normal((m2, n2))
    p = m2 if side == "L" else n2
    # entries scaled so that H's entries stay O(1) at large k
    V = rng.standard_normal((p, k) if storev == "C" else (k, p)) / np.sqrt(p)
    T = rng.standard_normal((k, k)) / np.sqrt(k)
    if dt == torch.float32:  # the oracle works on the values the device sees
        A1, A2, V, T = (x.astype(np.float32).astype(np.float64) for x in (A1, A2, V, T))
    work = np.zeros((k, n2) if side == "L" else (m2, k))
    e1, e2 = TO.parfb(side, trans, direct, storev, m1, n1, m2, n2, k, l, A1, A2, V, T, work)
    c1, c2 = TO.tprfb_dense(side, trans, direct, storev, l, V, T, A1, A2)
    assert rel(e1, c1) < 1e-12 and rel(e2, c2) < 1e-12
    A1d, A2d = dev(A1, dt, pad=3), dev(A2, dt, pad=5)
    parfb_(side, trans, direct, storev, A1d, A2d, dev(V, dt, pad=2), dev(T, dt, pad=1), k=k, l=l)
    torch.cuda.synchronize()
    assert rel(host(A1d), e1) < TOL[dt] and rel(host(A2d), e2) < TOL[dt]
    # the padding rows below each leading block are untouched
    for t, r in ((A1d, m1), (A2d, m2)):
        full = torch.as_strided(t, (t.stride(1), t.shape[1]), (1, t.stride(1)))
        assert torch.all(full[r:] == 7.25)


@pytest.mark.parametrize("dt", [torch.float64, torch.float32])
@pytest.mark.parametrize("op", ["W", "A"])
@pytest.mark.parametrize("side,storev,direct", [(s, v, d) for s in "LR" for v in "CR" for d in "FB"])
def test_pamm_ops_match_oracle(nla, dt, op, side, storev, direct):
    from nextla_amd import pamm_
    rng = np.random.default_rng(11)
    m, n, k, l = 40, 30, 20, 10  # test/pamm.jl:6
    A1 = rng.standard_normal((k, n) if side == "L" else (m, k))
    A2 = rng.standard_normal((m, n))
    p = m if side == "L" else n
    V = rng.standard_normal((p, k) if storev == "C" else (k, p)) / np.sqrt(p)
    W = rng.standard_normal((k, n) if side == "L" else (m, k))
    if dt == torch.float32:
        A1, A2, V, W = (x.astype(np.float32).astype(np.float64) for x in (A1, A2, V, W))
    eW, eA2 = TO.pamm(op, side, storev, direct, m, n, k, l, A1, A2, V, W)
    A1d, A2d, Vd, Wd = dev(A1, dt), dev(A2, dt), dev(V, dt), dev(W, dt)
    pamm_(op, side, storev, direct, m, n, k, l, A1d, A2d, Vd, Wd)
    torch.cuda.synchronize()
    assert rel(host(Wd), eW) < TOL[dt]
    assert rel(host(A2d), eA2) < TOL[dt]


def test_pamm_quick_returns_and_repeatability(nla):
    """test/pamm.jl:85-106 (m, n or k zero leave W as it was) and :146-162
    (two identical calls give identical W)."""
    from nextla_amd import pamm_
    rng = np.random.default_rng(5)
    A1 = dev(rng.standard_normal((5, 5)), torch.float64)
    A2 = dev(rng.standard_normal((10, 5)), torch.float64)
    V = dev(rng.standard_normal((10, 5)), torch.float64)
    W = dev(rng.standard_normal((5, 5)), torch.float64)
    W0 = W.clone()
    for m, n, k in ((0, 5, 5), (10, 0, 5), (10, 5, 0)):
        pamm_("W", "L", "C", "F", m, n, k, 3, A1, A2, V, W)
        torch.cuda.synchronize()
        assert torch.equal(W, W0)
    W1, W2 = torch.zeros_like(W), torch.zeros_like(W)
    pamm_("W", "L", "C", "F", 10, 5, 5, 3, A1, A2, V, W1)
    pamm_("W", "L", "C", "F", 10, 5, 5, 3, A1, A2, V, W2)
    torch.cuda.synchronize()
    assert torch.equal(W1, W2)


def test_argument_errors(nla):
    """test/pamm.jl:166-180, test/parfb.jl:55-66: bad characters and sizes are
    rejected before any device work, with the argument's position."""
    from nextla_amd import _lib
    lib = _lib.load()
    z = torch.zeros((16, 16), dtype=torch.float64, device=DEV)
    p = z.data_ptr()
    s = torch.cuda.current_stream().cuda_stream
    ch = _lib.ch
    assert lib.nla_pamm(ch("X"), ch("L"), ch("C"), ch("F"), 1, 10, 5, 5, 3, p, 5, p, 10, p, 10, p, 5, s) == -1
    assert lib.nla_pamm(ch("W"), ch("X"), ch("C"), ch("F"), 1, 10, 5, 5, 3, p, 5, p, 10, p, 10, p, 5, s) == -2
    assert lib.nla_pamm(ch("W"), ch("L"), ch("X"), ch("F"), 1, 10, 5, 5, 3, p, 5, p, 10, p, 10, p, 5, s) == -3
    assert lib.nla_pamm(ch("W"), ch("L"), ch("C"), ch("X"), 1, 10, 5, 5, 3, p, 5, p, 10, p, 10, p, 5, s) == -4
    assert lib.nla_pamm(ch("W"), ch("L"), ch("C"), ch("F"), 1, -1, 5, 5, 3, p, 5, p, 10, p, 10, p, 5, s) == -6
    assert lib.nla_pamm(ch("W"), ch("L"), ch("C"), ch("F"), 1, 10, -1, 5, 3, p, 5, p, 10, p, 10, p, 5, s) == -7
    assert lib.nla_pamm(ch("W"), ch("L"), ch("C"), ch("F"), 1, 10, 5, 5, 6, p, 5, p, 10, p, 10, p, 5, s) == -9
    args = (1, 4, 4, 4, 4, 2, 0, p, 2, p, 4, p, 4, p, 2, p, 2, s)
    assert lib.nla_parfb(ch("X"), ch("N"), ch("F"), ch("C"), *args) == -1
    assert lib.nla_parfb(ch("L"), ch("X"), ch("F"), ch("C"), *args) == -2
    assert lib.nla_parfb(ch("L"), ch("N"), ch("X"), ch("C"), *args) == -3
    assert lib.nla_parfb(ch("L"), ch("N"), ch("F"), ch("X"), *args) == -4
    # n1 != n2 on the left, m1 != m2 on the right, l > k, fewer than k rows of A1
    assert lib.nla_parfb(ch("L"), ch("N"), ch("F"), ch("C"), 1, 2, 3, 4, 4, 2, 0, p, 2, p, 4, p, 4, p, 2, p, 2,
                         s) == -9
    assert lib.nla_parfb(ch("R"), ch("N"), ch("F"), ch("C"), 1, 3, 2, 4, 4, 2, 0, p, 3, p, 4, p, 4, p, 2, p, 4,
                         s) == -8
    assert lib.nla_parfb(ch("L"), ch("N"), ch("F"), ch("C"), 1, 2, 4, 4, 4, 2, 3, p, 2, p, 4, p, 4, p, 2, p, 2,
                         s) == -11
    assert lib.nla_parfb(ch("L"), ch("N"), ch("F"), ch("C"), 1, 1, 4, 4, 4, 2, 0, p, 2, p, 4, p, 4, p, 2, p, 2,
                         s) == -6
    assert torch.all(z == 0)


def test_parfb_taller_a1_as_tsmqr_passes(nla):
    """tsmqr! (src/tsmqr.jl:132) hands parfb! an A1 block of m1 > k rows:
    only its first k rows change, as in the reference's loops over 1:k."""
    from nextla_amd import parfb_
    rng = np.random.default_rng(9)
    k, m1, m2, n2 = 8, 13, 20, 24
    A1 = rng.standard_normal((m1, n2))
    A2 = rng.standard_normal((m2, n2))
    V = rng.standard_normal((m2, k)) / np.sqrt(m2)
    T = rng.standard_normal((k, k)) / np.sqrt(k)
    e1, e2 = TO.parfb("L", "T", "F", "C", k, n2, m2, n2, k, 0, A1[:k], A2, V, T, np.zeros((k, n2)))
    A1d, A2d = dev(A1, torch.float64), dev(A2, torch.float64)
    parfb_("L", "T", "F", "C", A1d, A2d, dev(V, torch.float64), dev(T, torch.float64), k=k, l=0)
    torch.cuda.synchronize()
    out = host(A1d)
    assert rel(out[:k], e1) < 1e-12 and rel(host(A2d), e2) < 1e-12
    assert np.array_equal(out[k:], A1[k:])
