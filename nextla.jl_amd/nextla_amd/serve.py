"""Continuous-batching request server for triangular solves / multiplies on
one GPU (no counterpart in the reference, which serves one call at a time).

A request is one unified_rectrxm!(side, uplo, trans, alpha, func, A, B) call
(src/rectrxm.jl:43) whose A and B already live in HBM.  The server

  * admits requests against an HBM budget (the bytes of A, B and the
    batched workspace they need; A counted once while it is resident),
  * groups admitted requests that share (dtype, side, uplo, trans, func)
    into one batched plan (nla_batch_*: persistent panel launches over all
    (system, strip) tasks, systems above 1024 rows in 512-row stages with a
    grouped GEMM per stage), so a batch costs a fixed number of launches
    whatever its sizes,
  * keeps several batches in flight on separate streams, so requests that
    arrive while others run are admitted as slots and memory free up and
    their work overlaps on the device,
  * records each request's latency from arrival to completion (the
    completion event of its batch).

FP64 requests are FP64-accuracy requests.  Their TRSMs are served either by
the plain FP64 path (fp64_trsm="plain", batched like the rest) or by fp32
correction solves with FP64 refinement to backward error <= 1e-14
(fp64_trsm="mixed"): same-signature requests are batched into one
MixedBatchPlan (nla_batch_mixed_*: batched fp32 solves, batched FP64 TRMM
residuals, each system frozen once it reaches the target).  The refinement
reads its stopping test on the host once per step, so each refined batch
runs on a worker thread with its own stream while the loop keeps admitting
and launching.  The plain path is the faster of the two on MI355X (FP64 MFMA
runs at the FP64 vector rate, so each FP64 residual costs about as much as a
whole plain solve); "mixed" is there to serve and measure the refined form.

A system's result does not depend on what else is in its batch: the batched
kernel solves each (system, strip) task with the same code path whatever the
batch holds, so a request gives bit-identical output alone and in the mix
(a refined request likewise: its steps, error and result do not depend on
the rest of its batch).
"""
import collections
import concurrent.futures
import threading
import time

import torch

from . import unified_rectrxm_
from .batch import BATCH_N_MAX, BatchPlan, MixedBatchPlan


class Request:
    __slots__ = ("rid", "side", "uplo", "trans", "func", "alpha", "A", "B", "arrival", "done", "a_key", "ready")

    def __init__(self, rid, side, uplo, trans, func, alpha, A, B, arrival=None, a_key=None):
        self.rid = rid
        self.side, self.uplo, self.trans, self.func = side, uplo, trans, func
        self.alpha = alpha
        self.A, self.B = A, B
        self.arrival = arrival
        self.done = None
        self.a_key = a_key  # identifies a device-resident A shared between requests
        self.ready = None  # event on the submitting stream after A and B were written

    def wait_ready(self, stream):
        """Order `stream` after the submitter's writes of A and B, and tell the
        caching allocator that `stream` uses them (so memory the submitter
        frees is not reused while the server's work still reads it)."""
        if self.ready is not None:
            stream.wait_event(self.ready)
        self.A.record_stream(stream)
        self.B.record_stream(stream)

    def signature(self):
        return (self.A.dtype, self.side, self.uplo, self.trans, self.func)

    def nbytes(self):
        return self.B.numel() * self.B.element_size()


class _InFlight:
    __slots__ = ("reqs", "plan", "event", "stream", "nbytes", "a_keys", "future")


class _MixedWorkers:
    """Worker threads for refined FP64 batches, each with its own stream."""

    def __init__(self, device, workers, precision, tol, max_iter=10):
        self.device = device
        self.precision = precision
        self.tol = tol
        self.max_iter = max_iter
        self.local = threading.local()
        self.pool = concurrent.futures.ThreadPoolExecutor(max_workers=workers)

    def _stream(self):
        st = getattr(self.local, "stream", None)
        if st is None:
            st = self.local.stream = torch.cuda.Stream(device=self.device)
        return st

    def run(self, reqs):
        """Refine a batch of same-signature FP64 TRSM requests; returns the
        backward error of each."""
        from .mixed import rectrxm_mixed_
        stream = self._stream()
        for r in reqs:
            r.wait_ready(stream)
        small = [r for r in reqs if r.A.shape[0] <= BATCH_N_MAX]
        big = [r for r in reqs if r.A.shape[0] > BATCH_N_MAX]
        errors = {}
        with torch.cuda.stream(stream):
            if small:
                r0 = small[0]
                plan = MixedBatchPlan(r0.side, r0.uplo, r0.trans, [r.A for r in small], [r.B for r in small],
                                      [r.alpha for r in small], stream=stream)
                bwd, _, _ = plan.execute(tol=self.tol, max_iter=self.max_iter, stream=stream)
                plan.close()
                errors.update({r.rid: b for r, b in zip(small, bwd)})
            for r in big:  # beyond the batched planner's size limit: one refined call each
                _, info = rectrxm_mixed_(r.side, r.uplo, r.trans, r.alpha, r.A, r.B, precision=self.precision,
                                         tol=self.tol)
                errors[r.rid] = info["backward_error"]
        now = time.perf_counter()
        for r in reqs:
            r.done = now
        return errors

    def close(self):
        self.pool.shutdown(wait=True)


class Server:
    """Continuous-batching server.

    budget_bytes: HBM admission budget for in-flight requests.
    max_inflight: batches in flight at once (one stream each).
    max_batch: most requests per batch.
    use_graph: replay batched plans from hipGraphs.
    fp64_trsm: "plain" (FP64 TRSMs batched like the rest) or "mixed" (batched
        fp32 correction solves with FP64 refinement to mixed_tol on
        mixed_workers worker threads; mixed_precision is the update precision
        of the single-system refinement used above the batched size limit).
    direct_rows: requests with more rows go through the recursion one call
        each instead of into the batched plan (default: none; measured slower
        in aggregate, see the serving notes).
    """

    def __init__(self, device="cuda", budget_bytes=64 << 30, max_inflight=4, max_batch=512, use_graph=False,
                 fp64_trsm="plain", mixed_precision="fp16", mixed_tol=1e-14, mixed_workers=4, direct_rows=None):
        if fp64_trsm not in ("plain", "mixed"):
            raise ValueError(f"fp64_trsm must be 'plain' or 'mixed', got {fp64_trsm!r}")
        self.device = torch.device(device)
        self.budget = int(budget_bytes)
        self.max_batch = max_batch
        self.use_graph = use_graph
        self.streams = [torch.cuda.Stream(device=self.device) for _ in range(max_inflight)]
        self.free_streams = list(self.streams)
        self.queue = collections.deque()
        self.inflight = []
        self.used = 0
        self.resident = collections.OrderedDict()  # a_key -> (A tensor, refcount)
        self.completed = []
        self.fp64_trsm = fp64_trsm
        self.mixed = (_MixedWorkers(self.device, mixed_workers, mixed_precision, mixed_tol)
                      if fp64_trsm == "mixed" else None)
        self.mixed_slots = mixed_workers
        # requests with more rows than this go through the recursion one call
        # each (same stream, no wait) instead of into the batched plan
        self.direct_rows = BATCH_N_MAX if direct_rows is None else min(int(direct_rows), BATCH_N_MAX)

    # -- resident factors ----------------------------------------------------
    def resident_a(self, key, A=None):
        """Return the device-resident A stored under key (storing A first if
        given).  Resident factors are charged to the budget once."""
        if key in self.resident:
            self.resident.move_to_end(key)
            return self.resident[key]
        if A is None:
            raise KeyError(key)
        self.resident[key] = A
        self.used += A.numel() * A.element_size()
        return A

    def _evict(self, need):
        for key in list(self.resident):
            if self.used + need <= self.budget:
                break
            if any(r.a_key == key for r in self.queue) or any(
                    key in f.a_keys for f in self.inflight):
                continue
            A = self.resident.pop(key)
            self.used -= A.numel() * A.element_size()

    # -- request flow -------------------------------------------------------
    def submit(self, req):
        """Queue a request.  Its A and B may still be being written by work
        queued on the caller's current stream: the request's batch waits for
        that work (an event recorded here) before reading them."""
        if req.arrival is None:
            req.arrival = time.perf_counter()
        req.ready = torch.cuda.Event()
        req.ready.record(torch.cuda.current_stream(self.device))
        self.queue.append(req)

    def _is_mixed(self, r):
        return self.mixed is not None and r.func == "S" and r.A.dtype == torch.float64

    def _req_bytes(self, r):
        b = r.nbytes()
        if r.a_key is None:
            b += r.A.numel() * r.A.element_size()
        n = r.A.shape[0]
        if self._is_mixed(r):
            m = r.B.shape[1] if r.side == "L" else r.B.shape[0]
            b += 4 * n * n + 20 * n * m + 256 * n + 4096  # refinement arrays, fp32 inverses, descriptors
        elif r.func == "S":
            b += ((n + 63) // 64) * 64 * 64 * r.A.element_size()  # diagonal inverses
        return b

    def _launch_mixed(self):
        """Hand the queued refined requests that share the first one's
        signature (within the budget) to a free worker as one batch."""
        busy = sum(1 for f in self.inflight if f.future is not None)
        if busy >= self.mixed_slots:
            return False
        head = next((r for r in self.queue if self._is_mixed(r)), None)
        if head is None:
            return False
        sig = head.signature()
        take, keep, nbytes = [], collections.deque(), 0
        while self.queue and len(take) < self.max_batch:
            r = self.queue.popleft()
            rb = self._req_bytes(r)
            if self._is_mixed(r) and r.signature() == sig and self.used + nbytes + rb <= self.budget:
                take.append(r)
                nbytes += rb
            else:
                keep.append(r)
        self.queue.extendleft(reversed(keep))
        if not take:
            need = self._req_bytes(head)
            self._evict(need)
            if not self.inflight and self.used + need > self.budget:
                self.queue.remove(head)
                raise RuntimeError(f"request {head.rid} needs {need} bytes, more than the HBM budget of "
                                   f"{self.budget}")
            return False
        f = _InFlight()
        f.reqs, f.stream, f.nbytes, f.plan, f.event = take, None, nbytes, None, None
        f.a_keys = {r.a_key for r in take if r.a_key is not None}
        f.future = self.mixed.pool.submit(self.mixed.run, take)
        self.used += nbytes
        self.inflight.append(f)
        return True

    def _launch(self):
        if self.mixed is not None and self._launch_mixed():
            return True
        head = next((r for r in self.queue if not self._is_mixed(r)), None)
        if head is None or not self.free_streams:
            return False
        sig = head.signature()
        take, keep, nbytes = [], collections.deque(), 0
        while self.queue and len(take) < self.max_batch:
            r = self.queue.popleft()
            rb = self._req_bytes(r)
            if not self._is_mixed(r) and r.signature() == sig and self.used + nbytes + rb <= self.budget:
                take.append(r)
                nbytes += rb
            else:
                keep.append(r)
        self.queue.extendleft(reversed(keep))
        if not take:
            self._evict(self._req_bytes(head))
            if not self.inflight and self.used + self._req_bytes(head) > self.budget:
                self.queue.remove(head)
                raise RuntimeError(f"request {head.rid} needs {self._req_bytes(head)} bytes, "
                                   f"more than the HBM budget of {self.budget}")
            return False
        stream = self.free_streams.pop()
        f = _InFlight()
        f.reqs, f.stream, f.nbytes, f.future = take, stream, nbytes, None
        f.a_keys = {r.a_key for r in take if r.a_key is not None}
        small = [r for r in take if r.A.shape[0] <= self.direct_rows]
        big = [r for r in take if r.A.shape[0] > self.direct_rows]
        for r in take:
            r.wait_ready(stream)
        with torch.cuda.stream(stream):
            f.plan = None
            if small:
                dt, side, uplo, trans, func = sig
                f.plan = BatchPlan(side, uplo, trans, func, [r.A for r in small], [r.B for r in small],
                                   [r.alpha for r in small], stream=stream)
                f.plan.execute(use_graph=self.use_graph, stream=stream)
            for r in big:
                unified_rectrxm_(r.side, r.uplo, r.trans, r.alpha, r.func, r.A, r.B)
            f.event = torch.cuda.Event()
            f.event.record(stream)
        self.used += nbytes
        self.inflight.append(f)
        return True

    def _reap(self):
        done = [f for f in self.inflight if (f.future.done() if f.future is not None else f.event.query())]
        now = time.perf_counter()
        for f in done:
            self.inflight.remove(f)
            if f.future is not None:
                f.future.result()  # re-raises a worker's error here
                self.completed.extend(f.reqs)  # done was stamped by the worker
            else:
                for r in f.reqs:
                    r.done = now
                    self.completed.append(r)
                self.free_streams.append(f.stream)
            if f.plan is not None:
                f.plan.close()
            self.used -= f.nbytes
        return len(done)

    def step(self):
        """One turn of the loop: retire finished batches, launch new ones."""
        n = self._reap()
        launched = False
        while self._launch():
            launched = True
        if not n and not launched and self.mixed is not None:
            # Nothing to do but wait: block on the refinement workers for a
            # moment instead of polling, so their Python side gets the GIL.
            pending = [f.future for f in self.inflight if f.future is not None]
            if pending:
                concurrent.futures.wait(pending, timeout=1e-4, return_when=concurrent.futures.FIRST_COMPLETED)
        return n

    def drain(self):
        while self.queue or self.inflight:
            self.step()

    def idle(self):
        return not self.queue and not self.inflight

    def close(self):
        if self.mixed is not None:
            self.mixed.close()


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
