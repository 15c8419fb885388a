"""Continuous-batching request server for triangular solves / multiplies on
one GPU (no counterpart in the reference, which serves one call at a time):
a thin client of the server core behind the C ABI (nla_serve_* of
include/nla.h, csrc/nla_serve.cpp).

A request is one unified_rectrxm!(side, uplo, trans, alpha, func, A, B) call
(src/rectrxm.jl:43) whose A and B already live in HBM.  The core

  * admits requests against an HBM budget (the bytes of A, B and the
    batched workspace they need; A counted once while it is resident),
  * groups admitted requests that share (dtype, side, uplo, trans, func)
    into one batched plan (nla_batch_*: persistent panel launches over all
    (system, strip) tasks, systems above 1024 rows in 512-row stages with a
    grouped GEMM per stage), so a batch costs a fixed number of launches
    whatever its sizes,
  * keeps several batches in flight on separate streams, so requests that
    arrive while others run are admitted as streams and memory free up and
    their work overlaps on the device,
  * orders each request after the work queued on the submitter's stream
    (an event recorded at submit).

This client keeps each request's tensors alive until its completion is
reported, and stamps Request.done when it collects the completion (its
latency runs from arrival to that moment).

FP64 requests are FP64-accuracy requests.  Their TRSMs are served either by
the plain FP64 path (fp64_trsm="plain", batched like the rest) or by fp32
correction solves with FP64 refinement to backward error <= mixed_tol
(fp64_trsm="mixed"): same-signature requests are batched into one refined
batch (nla_batch_mixed_*), run by the core's worker threads, each with its
own stream, because the refinement reads its stopping test on the host once
per step.

A system's result does not depend on what else is in its batch: the batched
kernel solves each (system, strip) task with the same code path whatever the
batch holds, so a request gives bit-identical output alone and in the mix
(a refined request likewise: its steps, error and result do not depend on
the rest of its batch).
"""
import ctypes
import time

import torch

from . import _lib, unified_rectrxm_
from .batch import BATCH_N_MAX  # noqa: F401  (the core's batched size limit)


class Request:
    __slots__ = ("rid", "side", "uplo", "trans", "func", "alpha", "A", "B", "arrival", "done", "a_key",
                 "status", "backward_error", "iterations")

    def __init__(self, rid, side, uplo, trans, func, alpha, A, B, arrival=None, a_key=None):
        self.rid = rid
        self.side, self.uplo, self.trans, self.func = side, uplo, trans, func
        self.alpha = alpha
        self.A, self.B = A, B
        self.arrival = arrival
        self.done = None
        self.a_key = a_key  # identifies a device-resident A shared between requests
        self.status = None
        self.backward_error = None  # refined requests: the per-system backward error
        self.iterations = None

    def signature(self):
        return (self.A.dtype, self.side, self.uplo, self.trans, self.func)

    def nbytes(self):
        return self.B.numel() * self.B.element_size()


def _colmajor_ld(t, name, rid):
    r, c = t.shape
    if r > 1 and t.stride(0) != 1:
        raise ValueError(f"request {rid}: {name} must be column-major (stride (1, ld))")
    ld = t.stride(1) if c > 1 else max(r, 1)
    if c > 1 and ld < r:
        raise ValueError(f"request {rid}: {name} must be column-major (stride (1, ld))")
    return max(ld, 1)


_DT = {torch.float32: 0, torch.float64: 1}


class Server:
    """Continuous-batching server (the nla_serve_* core).

    budget_bytes: HBM admission budget for in-flight requests.
    max_inflight: batches in flight at once (one stream each).
    max_batch: most requests per batch.
    use_graph: replay batched plans from hipGraphs.
    fp64_trsm: "plain" (FP64 TRSMs batched like the rest) or "mixed" (batched
        fp32 correction solves with FP64 refinement to mixed_tol on
        mixed_workers worker threads; mixed_precision is the update precision
        of the single-system refinement used above the batched size limit).
    """

    def __init__(self, device="cuda", budget_bytes=64 << 30, max_inflight=4, max_batch=512, use_graph=False,
                 fp64_trsm="plain", mixed_precision="fp16", mixed_tol=1e-14, mixed_workers=4):
        from .mixed import PRECISIONS
        if fp64_trsm not in ("plain", "mixed"):
            raise ValueError(f"fp64_trsm must be 'plain' or 'mixed', got {fp64_trsm!r}")
        self.device = torch.device(device)
        if self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        self.fp64_trsm = fp64_trsm
        self._lib = _lib.load()
        self._srv = ctypes.c_void_p()
        mode = _lib.NLA_SERVE_REFINED if fp64_trsm == "mixed" else _lib.NLA_SERVE_PLAIN
        with torch.cuda.device(self.device):
            st = self._lib.nla_serve_create(int(budget_bytes), int(max_inflight), int(max_batch), mode,
                                            int(mixed_workers) if fp64_trsm == "mixed" else 0, 1 if use_graph else 0,
                                            ctypes.byref(self._srv))
        _lib.check(st, "nla_serve_create")
        if fp64_trsm == "mixed":
            _lib.check(self._lib.nla_serve_set_refinement(self._srv, float(mixed_tol), 10, PRECISIONS[mixed_precision]),
                       "nla_serve_set_refinement")
        self.tickets = {}  # ticket -> Request, until its completion is collected
        self.resident = {}  # key -> A (kept alive while the core holds it)
        self._key_ids = {}
        self._id_keys = {}
        self.completed = []
        self._done = (_lib.ServeDone * 1024)()
        self._evict_buf = (ctypes.c_int64 * 256)()

    # -- resident factors ----------------------------------------------------
    def _key_id(self, key):
        if key not in self._key_ids:
            kid = len(self._key_ids) + 1
            self._key_ids[key] = kid
            self._id_keys[kid] = key
        return self._key_ids[key]

    def resident_a(self, key, A=None):
        """Return the device-resident A stored under key (storing A first if
        given).  The core charges a resident factor to the budget once, from
        the first request that uses it."""
        if key in self.resident:
            return self.resident[key]
        if A is None:
            raise KeyError(key)
        self.resident[key] = A
        self._key_id(key)
        return A

    def _sync_evicted(self):
        n = ctypes.c_int64()
        while True:
            _lib.check(self._lib.nla_serve_evicted(self._srv, self._evict_buf, 256, ctypes.byref(n)),
                       "nla_serve_evicted")
            for i in range(n.value):
                self.resident.pop(self._id_keys.get(self._evict_buf[i]), None)
            if n.value < 256:
                break

    # -- request flow -------------------------------------------------------
    def submit(self, req):
        """Queue a request.  Its A and B may still be being written by work
        queued on the caller's current stream: the core records an event there
        and runs the request after it."""
        if req.arrival is None:
            req.arrival = time.perf_counter()
        A, B = req.A, req.B
        if A.dtype not in _DT or B.dtype != A.dtype:
            raise ValueError(f"request {req.rid}: A and B must share float32 or float64")
        if not (A.is_cuda and B.is_cuda and A.device == self.device and B.device == self.device):
            raise ValueError(f"request {req.rid}: A and B must live on {self.device}")
        n = A.shape[0]
        side = req.side.upper()
        r = _lib.ServeRequest()
        r.side, r.uplo, r.trans, r.func = (_lib.ch(c.upper()) for c in (req.side, req.uplo, req.trans, req.func))
        r.dtype = _DT[A.dtype]
        r.alpha = float(req.alpha)
        r.n = n
        r.m = B.shape[1] if side == "L" else B.shape[0]
        r.A, r.lda = A.data_ptr(), _colmajor_ld(A, "A", req.rid)
        r.B, r.ldb = B.data_ptr(), _colmajor_ld(B, "B", req.rid)
        r.a_key = self._key_id(req.a_key) if req.a_key is not None else 0
        if req.a_key is not None and req.a_key not in self.resident:
            self.resident[req.a_key] = A
        ticket = ctypes.c_int64()
        st = self._lib.nla_serve_submit(self._srv, ctypes.byref(r), torch.cuda.current_stream(self.device).cuda_stream,
                                        ctypes.byref(ticket))
        _lib.check(st, "nla_serve_submit")
        self.tickets[ticket.value] = req

    def step(self):
        """One turn of the loop: retire finished batches, launch new ones,
        collect completions (stamping Request.done).  Returns how many
        requests completed."""
        nd = ctypes.c_int64()
        st = self._lib.nla_serve_step(self._srv, self._done, len(self._done), ctypes.byref(nd))
        now = time.perf_counter()
        failed = []
        for i in range(nd.value):
            d = self._done[i]
            req = self.tickets.pop(d.ticket)
            req.done = now
            req.status = d.status
            if self.fp64_trsm == "mixed" and req.func.upper() == "S" and req.A.dtype == torch.float64:
                req.backward_error, req.iterations = d.backward_error, d.iterations
            self.completed.append(req)
            if d.status != 0:
                failed.append(req)
        _lib.check(st, "nla_serve_step")
        self._sync_evicted()
        if failed:
            r = failed[0]
            raise RuntimeError(f"request {r.rid} failed with status {r.status}"
                               + (" (it needs more than the HBM budget)" if r.status == _lib.NLA_ERR_WORKSPACE else ""))
        return nd.value

    def drain(self):
        while self.tickets:
            self.step()

    def idle(self):
        return not self.tickets

    def _stats(self):
        s = _lib.ServeStats()
        _lib.check(self._lib.nla_serve_stats(self._srv, ctypes.byref(s)), "nla_serve_stats")
        return s

    @property
    def used(self):
        return self._stats().used_bytes

    @property
    def budget(self):
        return self._stats().budget_bytes

    def close(self):
        if getattr(self, "_srv", None) is not None and self._srv.value:
            st = self._lib.nla_serve_destroy(self._srv)
            self._srv = ctypes.c_void_p()
            _lib.check(st, "nla_serve_destroy")

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def workload_requests(count, seed=0, device="cuda", reuse=0.2, pool=8, n_range=(32, 4096), m_range=(1, 512),
                      cond_max=1e4):
    """The serving workload: count left-lower requests, TRSM / TRMM and FP32 /
    FP64 accuracy half and half, n log-uniform on n_range, right-hand sides
    log-uniform on m_range, A as the refinement workload (off-diagonal
    U(-1,1)/sqrt(n), log-spaced diagonal 2 .. 2c with c log-uniform on
    [1, cond_max / 2], so cond(A) <= cond_max), B N(0,1); a fraction `reuse`
    of the requests use one of `pool` device-resident factors per dtype
    (a_key set).  All inputs are generated in HBM and synchronized."""
    import math

    import numpy as np

    from .inputs import device_precision_tri
    rng = np.random.default_rng(seed)
    lg = math.log
    factors = {}
    reqs = []
    for k in range(count):
        dt = torch.float64 if rng.random() < 0.5 else torch.float32
        func = "S" if rng.random() < 0.5 else "M"
        m = int(math.exp(rng.uniform(lg(m_range[0]), lg(m_range[1]))))
        c = math.exp(rng.uniform(0.0, lg(cond_max / 2.0)))
        if rng.random() < reuse:
            key = (int(rng.integers(pool)), str(dt))
            if key not in factors:
                n = int(math.exp(rng.uniform(lg(n_range[0]), lg(n_range[1]))))
                A, _ = device_precision_tri(n, 1, c, seed * 100003 + len(factors), device)
                factors[key] = A if dt == torch.float64 else A.float()
            A, a_key = factors[key], key
            n = A.shape[0]
        else:
            n = int(math.exp(rng.uniform(lg(n_range[0]), lg(n_range[1]))))
            A, _ = device_precision_tri(n, 1, c, seed * 100003 + 50000 + k, device)
            A = A if dt == torch.float64 else A.float()
            a_key = None
        g = torch.Generator(device=device)
        g.manual_seed(seed * 7919 + k)
        B = torch.empty((m, n), dtype=dt, device=device).t().normal_(generator=g)
        reqs.append(Request(k, "L", "L", "N", func, 1.0, A, B, a_key=a_key))
    torch.cuda.synchronize()
    return reqs


def one_at_a_time(requests, arrivals=None):
    """Baseline: each request as its own unified_rectrxm! call, in arrival
    order, waiting for it before the next (one call per system, as the
    reference serves).  Returns per-request latencies in seconds."""
    lat = []
    t0 = time.perf_counter()
    for k, r in enumerate(requests):
        if arrivals is not None:
            target = t0 + arrivals[k]
            while time.perf_counter() < target:
                pass
            start = target
        else:
            start = time.perf_counter()
        unified_rectrxm_(r.side, r.uplo, r.trans, r.alpha, r.func, r.A, r.B)
        torch.cuda.current_stream().synchronize()
        lat.append(time.perf_counter() - start)
    return lat


def run_stream(server, requests, arrivals):
    """Feed requests at their arrival offsets (seconds from now) and run the
    server loop until all are done.  Returns (latencies in seconds, wall)."""
    t0 = time.perf_counter()
    k = 0
    n = len(requests)
    while k < n or not server.idle():
        now = time.perf_counter()
        while k < n and t0 + arrivals[k] <= now:
            requests[k].arrival = t0 + arrivals[k]
            server.submit(requests[k])
            k += 1
        server.step()
    wall = time.perf_counter() - t0
    return [r.done - r.arrival for r in requests], wall
