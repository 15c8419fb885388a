"""Batched triangular solves / multiplies of many independent systems on one
GPU (the nla_batch_* entry points of include/nla.h).

The reference solves each system with its own unified_rectrxm! call
(src/rectrxm.jl:43); here a whole list of systems, each with its own n, m,
A, B and alpha but a shared side / uplo / trans / func and dtype, runs in a
fixed number of launches: the batched inversion of the 64x64 diagonal blocks
(TRSM), then per stage a persistent panel launch per size class and one
grouped GEMM (systems up to 1024 rows are a single panel; larger ones are
solved in 512-row stages), optionally replayed from a hipGraph.

MixedBatchPlan is the FP64-accurate form of a TRSM batch: fp32 batched
correction solves and FP64 batched residuals until every system reaches the
requested backward error.
"""
import ctypes

import torch

from . import _lib
from ._lib import BatchItem

BATCH_MAX = 1024     # largest system solved as one panel
BATCH_N_MAX = 65536  # largest system a plan takes


def _items(side, As, Bs, alphas, dtypes):
    """nla_batch_item array for the (A_i, B_i) pairs; returns (items, dtype,
    device)."""
    if len(As) != len(Bs):
        raise ValueError("As and Bs differ in length")
    if not isinstance(alphas, (list, tuple)):
        alphas = [alphas] * len(As)
    dt = As[0].dtype if As else dtypes[-1]
    if dt not in dtypes:
        raise ValueError(f"unsupported element type {dt}")
    device = As[0].device if As else torch.device("cuda")
    count = len(As)
    items = (BatchItem * max(count, 1))()
    for i, (A, B) in enumerate(zip(As, Bs)):
        if A.dtype != dt or B.dtype != dt:
            raise ValueError("all A_i and B_i must share one element type")
        if not (A.is_cuda and B.is_cuda):
            raise ValueError("A_i and B_i must live on the GPU")
        n = A.shape[0]
        if A.dim() != 2 or A.shape[1] != n or B.dim() != 2:
            raise ValueError(f"system {i}: A must be square and B a matrix")
        if (side == "L" and B.shape[0] != n) or (side == "R" and B.shape[1] != n):
            raise ValueError(f"system {i}: B {tuple(B.shape)} does not match A for side {side!r}")
        lds = []
        for t, nm in ((A, "A"), (B, "B")):
            r, c = t.shape
            if r > 1 and t.stride(0) != 1:
                raise ValueError(f"system {i}: {nm} must be column-major (stride (1, ld))")
            ld = t.stride(1) if c > 1 else max(r, 1)
            if c > 1 and ld < r:
                raise ValueError(f"system {i}: {nm} must be column-major (stride (1, ld))")
            lds.append(max(ld, 1))
        m = B.shape[1] if side == "L" else B.shape[0]
        items[i].n, items[i].m = n, m
        items[i].A, items[i].lda = A.data_ptr(), lds[0]
        items[i].B, items[i].ldb = B.data_ptr(), lds[1]
        items[i].alpha = float(alphas[i])
    return items, dt, device


class BatchPlan:
    """A planned batch: descriptors and task queue resident in a device
    workspace owned by this object.  The A_i / B_i tensors must stay alive
    (the plan keeps their addresses); execute() may be called repeatedly,
    e.g. after refilling the B_i in place."""

    def __init__(self, side, uplo, trans, func, As, Bs, alphas=1.0, stream=None):
        side, uplo, trans, func = side.upper(), uplo.upper(), trans.upper(), func.upper()
        items, dt, self.device = _items(side, As, Bs, alphas, (torch.float32, torch.float64))
        self.dtype = dt
        count = len(As)
        lib = _lib.load()
        code = _lib.NLA_F64 if dt == torch.float64 else _lib.NLA_F32
        nbytes = lib.nla_batch_workspace_bytes(_lib.ch(func), code, items, count)
        if nbytes < 0:
            raise ValueError("invalid batch description")
        self.workspace = torch.empty(max(nbytes, 256), dtype=torch.uint8, device=self.device)
        self._plan = ctypes.c_void_p()
        self._keep = (list(As), list(Bs))
        s = (stream or torch.cuda.current_stream(self.device)).cuda_stream
        st = lib.nla_batch_plan_create(_lib.ch(side), _lib.ch(uplo), _lib.ch(trans), _lib.ch(func), code, items,
                                       count, self.workspace.data_ptr(), nbytes, s, ctypes.byref(self._plan))
        _lib.check(st, "nla_batch_plan_create")
        self.count = count

    def execute(self, use_graph=False, stream=None):
        """Enqueue the whole batch on `stream` (torch's current stream by
        default); with use_graph the launch sequence is captured into a
        hipGraph on the first call and replayed afterwards."""
        s = (stream or torch.cuda.current_stream(self.device)).cuda_stream
        st = _lib.load().nla_batch_plan_execute(self._plan, s, 1 if use_graph else 0)
        _lib.check(st, "nla_batch_plan_execute")

    def close(self):
        if getattr(self, "_plan", None) is not None and self._plan.value:
            _lib.load().nla_batch_plan_destroy(self._plan)
            self._plan = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batched_rectrxm_(side, uplo, trans, func, As, Bs, alphas=1.0, use_graph=False):
    """unified_rectrxm! applied to every (A_i, B_i) pair in one batched
    launch; returns the list of B_i (overwritten)."""
    plan = BatchPlan(side, uplo, trans, func, As, Bs, alphas)
    plan.execute(use_graph=use_graph)
    torch.cuda.current_stream(plan.device).synchronize()
    plan.close()
    return Bs


class MixedBatchPlan:
    """Batched FP64-accurate TRSM (nla_batch_mixed_*): fp32 correction solves
    of every system in one batched plan, FP64 residuals from one batched
    TRMM, each system refined until its normwise backward error is <= tol
    (then frozen).  A_i, B_i: float64 column-major device tensors that must
    stay alive while the plan exists; X_i overwrites B_i."""

    def __init__(self, side, uplo, trans, As, Bs, alphas=1.0, stream=None):
        side, uplo, trans = side.upper(), uplo.upper(), trans.upper()
        items, _, self.device = _items(side, As, Bs, alphas, (torch.float64,))
        count = len(As)
        lib = _lib.load()
        nbytes = lib.nla_batch_mixed_workspace_bytes(_lib.ch(side), items, count)
        if nbytes < 0:
            raise ValueError("invalid batch description")
        s = (stream or torch.cuda.current_stream(self.device))
        with torch.cuda.stream(s):
            self.workspace = torch.empty(max(nbytes, 256) + 256, dtype=torch.uint8, device=self.device)
        base = self.workspace.data_ptr()
        base += (-base) % 256
        self._plan = ctypes.c_void_p()
        self._keep = (list(As), list(Bs))
        st = lib.nla_batch_mixed_plan_create(_lib.ch(side), _lib.ch(uplo), _lib.ch(trans), items, count, base,
                                             nbytes, s.cuda_stream, ctypes.byref(self._plan))
        _lib.check(st, "nla_batch_mixed_plan_create")
        self.count = count
        self.workspace_bytes = nbytes

    def execute(self, tol=1e-14, max_iter=10, use_graph=False, stream=None, raise_on_stall=True):
        """Refine every system; returns (backward errors, corrections applied,
        converged) per system.  Synchronizes `stream` once per step."""
        n = max(self.count, 1)
        bwd = (ctypes.c_double * n)()
        its = (ctypes.c_int32 * n)()
        s = (stream or torch.cuda.current_stream(self.device)).cuda_stream
        st = _lib.load().nla_batch_mixed_plan_execute(self._plan, float(tol), int(max_iter), s,
                                                      1 if use_graph else 0, bwd, its)
        out_b = [bwd[i] for i in range(self.count)]
        out_i = [its[i] for i in range(self.count)]
        conv = [b <= tol for b in out_b]
        if st == _lib.NLA_ERR_NOT_CONVERGED and not raise_on_stall:
            return out_b, out_i, conv
        _lib.check(st, "nla_batch_mixed_plan_execute")
        return out_b, out_i, conv

    def close(self):
        if getattr(self, "_plan", None) is not None and self._plan.value:
            _lib.load().nla_batch_mixed_plan_destroy(self._plan)
            self._plan = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batched_rectrxm_mixed_(side, uplo, trans, As, Bs, alphas=1.0, tol=1e-14, max_iter=10, use_graph=False):
    """FP64-accurate unified_rectrxm! TRSM of every (A_i, B_i) pair by
    batched refinement; returns (Bs, backward errors, corrections)."""
    plan = MixedBatchPlan(side, uplo, trans, As, Bs, alphas)
    bwd, its, _ = plan.execute(tol=tol, max_iter=max_iter, use_graph=use_graph)
    plan.close()
    return Bs, bwd, its
