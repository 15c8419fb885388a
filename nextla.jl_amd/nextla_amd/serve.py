"""Continuous-batching request server for triangular solves / multiplies on
one GPU (no counterpart in the reference, which serves one call at a time).

A request is one unified_rectrxm!(side, uplo, trans, alpha, func, A, B) call
(src/rectrxm.jl:43) whose A and B already live in HBM.  The server

  * admits requests against an HBM budget (the bytes of A, B and the
    batched workspace they need; A counted once while it is resident),
  * groups admitted requests that share (dtype, side, uplo, trans, func)
    into one batched launch (nla_batch_*: a persistent kernel over all
    (system, strip) tasks) and sends systems larger than the batched
    solver's limit through the recursive path on the same stream,
  * keeps several batches in flight on separate streams, so requests that
    arrive while others run are admitted as slots and memory free up and
    their work overlaps on the device,
  * records each request's latency from arrival to completion (the
    completion event of its batch).

A system's result does not depend on what else is in its batch: the batched
kernel solves each (system, strip) task with the same code path whatever the
batch holds, so a request gives bit-identical output alone and in the mix.
"""
import collections
import time

import torch

from . import unified_rectrxm_
from .batch import BATCH_MAX, BatchPlan


class Request:
    __slots__ = ("rid", "side", "uplo", "trans", "func", "alpha", "A", "B", "arrival", "done", "a_key")

    def __init__(self, rid, side, uplo, trans, func, alpha, A, B, arrival=None, a_key=None):
        self.rid = rid
        self.side, self.uplo, self.trans, self.func = side, uplo, trans, func
        self.alpha = alpha
        self.A, self.B = A, B
        self.arrival = arrival
        self.done = None
        self.a_key = a_key  # identifies a device-resident A shared between requests

    def signature(self):
        return (self.A.dtype, self.side, self.uplo, self.trans, self.func)

    def nbytes(self):
        return self.B.numel() * self.B.element_size()


class _InFlight:
    __slots__ = ("reqs", "plan", "event", "stream", "nbytes", "a_keys")


class Server:
    """Continuous-batching server.

    budget_bytes: HBM admission budget for in-flight requests.
    max_inflight: batches in flight at once (one stream each).
    max_batch: most requests per batch.
    """

    def __init__(self, device="cuda", budget_bytes=64 << 30, max_inflight=4, max_batch=512, use_graph=False):
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
        if req.arrival is None:
            req.arrival = time.perf_counter()
        self.queue.append(req)

    def _req_bytes(self, r):
        b = r.nbytes()
        if r.a_key is None:
            b += r.A.numel() * r.A.element_size()
        if r.func == "S" and r.A.shape[0] <= BATCH_MAX:
            b += ((r.A.shape[0] + 63) // 64) * 64 * 64 * r.A.element_size()
        return b

    def _launch(self):
        if not self.queue or not self.free_streams:
            return False
        sig = self.queue[0].signature()
        take, keep, nbytes = [], collections.deque(), 0
        while self.queue and len(take) < self.max_batch:
            r = self.queue.popleft()
            rb = self._req_bytes(r)
            if r.signature() == sig and self.used + nbytes + rb <= self.budget:
                take.append(r)
                nbytes += rb
            else:
                keep.append(r)
        self.queue.extendleft(reversed(keep))
        if not take:
            head = self.queue[0]
            self._evict(self._req_bytes(head))
            if not self.inflight and self.used + self._req_bytes(head) > self.budget:
                self.queue.popleft()
                raise RuntimeError(f"request {head.rid} needs {self._req_bytes(head)} bytes, "
                                   f"more than the HBM budget of {self.budget}")
            return False
        stream = self.free_streams.pop()
        f = _InFlight()
        f.reqs, f.stream, f.nbytes = take, stream, nbytes
        f.a_keys = {r.a_key for r in take if r.a_key is not None}
        small = [r for r in take if r.A.shape[0] <= BATCH_MAX]
        big = [r for r in take if r.A.shape[0] > BATCH_MAX]
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
        done = [f for f in self.inflight if f.event.query()]
        now = time.perf_counter()
        for f in done:
            self.inflight.remove(f)
            for r in f.reqs:
                r.done = now
                self.completed.append(r)
            if f.plan is not None:
                f.plan.close()
            self.used -= f.nbytes
            self.free_streams.append(f.stream)
        return len(done)

    def step(self):
        """One turn of the loop: retire finished batches, launch new ones."""
        n = self._reap()
        while self._launch():
            pass
        return n

    def drain(self):
        while self.queue or self.inflight:
            self.step()

    def idle(self):
        return not self.queue and not self.inflight


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
