"""Continuous-batching server: every request matches a single unified_rectrxm!
call, a request's result is bit-identical alone and in a mixed stream, the
HBM budget is honoured, and resident factors are reused."""
import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu
DEV = "cuda"


def colmajor_rand(rows, cols, dt, gen):
    t = torch.empty((cols, rows), dtype=dt, device=DEV).t()
    t.normal_(generator=gen)
    return t


def tri(n, dt, gen):
    A = torch.empty((n, n), dtype=dt, device=DEV).t()
    A.uniform_(-1, 1, generator=gen).mul_(1.0 / n)
    A.diagonal().uniform_(1, 2, generator=gen)
    return A


def make_requests(nla, count, seed):
    from nextla_amd.serve import Request
    gen = torch.Generator(device=DEV)
    gen.manual_seed(seed)
    rng = np.random.default_rng(seed)
    reqs = []
    for k in range(count):
        n = int(np.exp(rng.uniform(np.log(32), np.log(1500))))
        m = int(np.exp(rng.uniform(np.log(1), np.log(200))))
        dt = torch.float64 if k % 2 == 0 else torch.float32
        func = "S" if k % 3 else "M"
        reqs.append(Request(k, "L", "L", "N", func, 1.0, tri(n, dt, gen), colmajor_rand(n, m, dt, gen)))
    return reqs


def test_server_matches_single_calls_and_alone(nla):
    from nextla_amd.serve import Request, Server
    reqs = make_requests(nla, 40, 1)
    B0 = [r.B.clone() for r in reqs]
    # expected: one plain call per request
    expect = []
    for r, b0 in zip(reqs, B0):
        x = b0.clone()
        nla.unified_rectrxm_(r.side, r.uplo, r.trans, r.alpha, r.func, r.A, x)
        expect.append(x)
    torch.cuda.synchronize()
    srv = Server(budget_bytes=8 << 30, max_inflight=3, max_batch=16)
    for r in reqs:
        srv.submit(r)
    srv.drain()
    torch.cuda.synchronize()
    eps = {torch.float64: 2.2e-16, torch.float32: 1.2e-7}
    for r, x in zip(reqs, expect):
        n = r.A.shape[0]
        rel = (torch.linalg.norm(r.B.double() - x.double()) / torch.linalg.norm(x.double())).item()
        assert rel <= 100 * n * eps[r.A.dtype], (r.rid, n, rel)
    # the same request alone gives bit-identical output
    for k in (0, 7, 13):
        r = reqs[k]
        solo = Request(1000 + k, r.side, r.uplo, r.trans, r.func, r.alpha, r.A, B0[k].clone())
        s2 = Server(budget_bytes=8 << 30, max_inflight=1, max_batch=1)
        s2.submit(solo)
        s2.drain()
        torch.cuda.synchronize()
        assert torch.equal(solo.B, r.B), k


def test_budget_forces_serial_admission(nla):
    from nextla_amd.serve import Server
    reqs = make_requests(nla, 12, 2)
    biggest = max(r.A.numel() * r.A.element_size() + r.B.numel() * r.B.element_size() + 64 * 64 * 8 * 32
                  for r in reqs)
    srv = Server(budget_bytes=biggest, max_inflight=4, max_batch=8)
    for r in reqs:
        srv.submit(r)
    peak = 0
    while not srv.idle():
        srv.step()
        peak = max(peak, srv.used)
        assert srv.used <= srv.budget
    assert all(r.done is not None for r in reqs)


def test_resident_factor_reuse(nla):
    from nextla_amd.serve import Request, Server
    gen = torch.Generator(device=DEV)
    gen.manual_seed(3)
    srv = Server(budget_bytes=4 << 30)
    A = srv.resident_a("factor-0", tri(300, torch.float64, gen))
    reqs = [Request(k, "L", "L", "N", "S", 1.0, srv.resident_a("factor-0"), colmajor_rand(300, 8, torch.float64, gen),
                    a_key="factor-0") for k in range(6)]
    B0 = [r.B.clone() for r in reqs]
    for r in reqs:
        srv.submit(r)
    srv.drain()
    torch.cuda.synchronize()
    Ah = A.cpu().numpy()
    for r, b0 in zip(reqs, B0):
        e = O.backward_error("L", "L", "N", 1.0, Ah, b0.cpu().numpy(), r.B.cpu().numpy())
        assert e < 10 * 300 * 2.2e-16
    assert "factor-0" in srv.resident


def test_refined_fp64_requests(nla):
    """fp64_trsm='mixed': FP64 TRSM requests are served by refinement to
    backward error <= 1e-14, agree with the plain FP64 solve, and give the
    same bits alone and in the mixed stream."""
    from nextla_amd import serve
    reqs = serve.workload_requests(24, seed=5, n_range=(32, 1500), m_range=(1, 64))
    B0 = [r.B.clone() for r in reqs]
    srv = serve.Server(budget_bytes=16 << 30, max_inflight=3, max_batch=8, fp64_trsm="mixed", mixed_workers=2)
    for r in reqs:
        srv.submit(r)
    srv.drain()
    srv.close()
    torch.cuda.synchronize()
    refined = [k for k, r in enumerate(reqs) if r.func == "S" and r.A.dtype == torch.float64]
    assert refined
    for k in refined:
        r, b0 = reqs[k], B0[k]
        T = torch.tril(r.A)
        bwd = (torch.linalg.norm(b0 - T @ r.B) / (torch.linalg.norm(T) * torch.linalg.norm(r.B)
                                                  + torch.linalg.norm(b0))).item()
        assert bwd <= 1.5e-14, (k, bwd)
        P = b0.clone()
        nla.unified_rectrxm_("L", "L", "N", 1.0, "S", r.A, P)
        rel = (torch.linalg.norm(r.B - P) / torch.linalg.norm(P)).item()
        assert rel <= 1e-10, (k, rel)
    k = refined[0]
    solo = serve.Request(999, "L", "L", "N", "S", 1.0, reqs[k].A, B0[k].clone())
    s2 = serve.Server(budget_bytes=16 << 30, fp64_trsm="mixed", mixed_workers=1)
    s2.submit(solo)
    s2.drain()
    s2.close()
    torch.cuda.synchronize()
    assert torch.equal(solo.B, reqs[k].B)


COMBOS8 = [(s, u, t) for s in "LR" for u in "LU" for t in "NT"]


@pytest.mark.parametrize("dt", [torch.float64, torch.float32])
def test_every_signature_poisson_stream(nla, dt):
    """Every side/uplo/trans x TRSM/TRMM, n up to 4096, m up to 512, fed at
    Poisson arrivals (2,000 requests/s): each result agrees with its single
    unified_rectrxm! call."""
    from nextla_amd import serve
    gen = torch.Generator(device=DEV)
    gen.manual_seed(21)
    rng = np.random.default_rng(21)
    reqs, expect = [], []
    for k in range(48):
        side, uplo, trans = COMBOS8[k % 8]
        func = "S" if (k // 8) % 2 == 0 else "M"
        n = int(np.exp(rng.uniform(np.log(32), np.log(4096))))
        m = int(np.exp(rng.uniform(np.log(1), np.log(512))))
        A = tri(n, dt, gen)
        B = colmajor_rand(n, m, dt, gen) if side == "L" else colmajor_rand(m, n, dt, gen)
        x = B.clone()
        nla.unified_rectrxm_(side, uplo, trans, 0.5, func, A, x)
        expect.append(x)
        reqs.append(serve.Request(k, side, uplo, trans, func, 0.5, A, B))
    torch.cuda.synchronize()
    arrivals = np.cumsum(rng.exponential(1.0 / 2000.0, size=len(reqs)))
    srv = serve.Server(budget_bytes=16 << 30, max_inflight=4, max_batch=32)
    lat, wall = serve.run_stream(srv, reqs, arrivals)
    srv.close()
    torch.cuda.synchronize()
    eps = {torch.float64: 2.2e-16, torch.float32: 1.2e-7}[dt]
    for r, x in zip(reqs, expect):
        n = r.A.shape[0]
        rel = (torch.linalg.norm(r.B.double() - x.double()) / torch.linalg.norm(x.double())).item()
        assert rel <= 100 * n * eps, (r.rid, r.side, r.uplo, r.trans, r.func, n, rel)
        assert r.status == 0
    assert all(v >= 0 for v in lat)


def test_c_abi_server_core(nla):
    """nla_serve_* called directly: tickets, completion records, a resident
    factor submitted with A = NULL after its first use, a request larger than
    the budget failed with NLA_ERR_WORKSPACE (the others unaffected), stats."""
    import ctypes
    from nextla_amd import _lib
    lib = _lib.load()
    gen = torch.Generator(device=DEV)
    gen.manual_seed(4)
    A = tri(200, torch.float64, gen)
    Bs = [colmajor_rand(200, 5, torch.float64, gen) for _ in range(3)]
    B0 = [b.clone() for b in Bs]
    budget = 8 << 20
    srv = ctypes.c_void_p()
    assert lib.nla_serve_create(budget, 2, 8, _lib.NLA_SERVE_PLAIN, 0, 0, ctypes.byref(srv)) == 0
    s = torch.cuda.current_stream().cuda_stream
    tickets = []
    for k, B in enumerate(Bs):
        r = _lib.ServeRequest(b"L", b"L", b"N", b"S", 1, 1.0, 200, 5, A.data_ptr() if k == 0 else None, 200,
                              B.data_ptr(), 200, 7)
        t = ctypes.c_int64()
        assert lib.nla_serve_submit(srv, ctypes.byref(r), s, ctypes.byref(t)) == 0
        tickets.append(t.value)
    big = torch.zeros((2048, 2048), dtype=torch.float64, device=DEV)  # 32 MB of A: above the budget
    bigB = torch.zeros((2048, 1), dtype=torch.float64, device=DEV)
    r = _lib.ServeRequest(b"L", b"L", b"N", b"M", 1, 1.0, 2048, 1, big.data_ptr(), 2048, bigB.data_ptr(), 2048, 0)
    t = ctypes.c_int64()
    assert lib.nla_serve_submit(srv, ctypes.byref(r), s, ctypes.byref(t)) == 0
    big_ticket = t.value
    assert lib.nla_serve_drain(srv) == 0
    st = _lib.ServeStats()
    assert lib.nla_serve_stats(srv, ctypes.byref(st)) == 0
    assert st.queued == 0 and st.running == 0 and st.completed_pending == 4 and st.resident == 1
    assert st.used_bytes == 200 * 200 * 8 and st.budget_bytes == budget  # only the resident factor stays
    done = (_lib.ServeDone * 8)()
    nd = ctypes.c_int64()
    assert lib.nla_serve_step(srv, done, 8, ctypes.byref(nd)) == 0 and nd.value == 4
    status = {done[i].ticket: done[i].status for i in range(nd.value)}
    assert status[big_ticket] == _lib.NLA_ERR_WORKSPACE
    assert all(status[t] == 0 for t in tickets)
    keys = (ctypes.c_int64 * 4)()
    cnt = ctypes.c_int64()
    assert lib.nla_serve_evicted(srv, keys, 4, ctypes.byref(cnt)) == 0 and cnt.value == 0
    # a request that fits only once the unused resident factor is evicted
    A2 = tri(980, torch.float64, gen)
    B2 = colmajor_rand(980, 1, torch.float64, gen)
    B20 = B2.clone()
    r = _lib.ServeRequest(b"L", b"L", b"N", b"S", 1, 1.0, 980, 1, A2.data_ptr(), 980, B2.data_ptr(), 980, 0)
    assert lib.nla_serve_submit(srv, ctypes.byref(r), s, ctypes.byref(t)) == 0
    assert lib.nla_serve_drain(srv) == 0
    assert lib.nla_serve_step(srv, done, 8, ctypes.byref(nd)) == 0 and nd.value == 1 and done[0].status == 0
    assert lib.nla_serve_evicted(srv, keys, 4, ctypes.byref(cnt)) == 0 and cnt.value == 1 and keys[0] == 7
    assert lib.nla_serve_stats(srv, ctypes.byref(st)) == 0 and st.resident == 0 and st.used_bytes == 0
    assert lib.nla_serve_destroy(srv) == 0
    x = B20.clone()
    nla.unified_rectrxm_("L", "L", "N", 1.0, "S", A2, x)
    assert (torch.linalg.norm(B2 - x) / torch.linalg.norm(x)).item() < 1e-13
    for B, b0 in zip(Bs, B0):
        x = b0.clone()
        nla.unified_rectrxm_("L", "L", "N", 1.0, "S", A, x)
        assert (torch.linalg.norm(B - x) / torch.linalg.norm(x)).item() < 1e-13
