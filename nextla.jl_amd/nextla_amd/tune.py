"""Autotuner for the shapes below the recursion and of the batched solver.

The reference fixes its leaf thresholds (256 for TRSM, 16 for TRMM,
src/rectrxm.jl:52,63); here the corresponding choices are measured on the
device at hand and kept per device:

  panel rows / strip  the largest diagonal block one fused panel launch
                      solves (a multiple of 64 up to 1024; 256-512 are
                      measured) and the right-hand
                      sides per work-group (16 or 32), per dtype
                      (NLA_CFG_PANEL_ROWS_*, NLA_CFG_PANEL_STRIP_*); a
                      tuned size replaces the library's default (0: 512
                      rows where the panel's rows are contiguous in
                      memory, 256 where they are strided) for every
                      layout, so it is tuned on the layout that matters;
  batch classes       which LDS-sized variants of the persistent batched
                      solver a plan may use, and whether systems above 512
                      rows are one 1024-row panel or two 512-row stages
                      (NLA_CFG_BATCH_CLASSES).

Each candidate is timed with HIP events on a representative problem (the
8192 x 8192 drop-in case for panels; a 10,000-system-style mixed batch for
the classes) and the fastest is applied with nla_set_config.  Results are
cached as JSON per device name (NEXTLA_AMD_TUNE_CACHE, else
~/.cache/nextla_amd/tune.json); tune() reuses a cached entry unless told
to measure again.  Every candidate gives results identical to the
oracle's within the same bounds (tests/test_gpu_parity.py::test_panel_shapes,
tests/test_gpu_batch.py::test_size_classes_bitwise_identical), so the tuner
trades speed only.
"""
import json
import math
import os

from . import _lib

PANEL_ROWS = (256, 384, 512)
PANEL_STRIPS = (16, 32)
BATCH_MASKS = (0x8, 0xC, 0xE, 0xF, 0x1E, 0x1F)  # 0x10: stage systems above 512 rows
_KEYS = {"f32": (_lib.NLA_CFG_PANEL_ROWS_F32, _lib.NLA_CFG_PANEL_STRIP_F32),
         "f64": (_lib.NLA_CFG_PANEL_ROWS_F64, _lib.NLA_CFG_PANEL_STRIP_F64)}


def cache_path():
    return os.environ.get("NEXTLA_AMD_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "nextla_amd", "tune.json"))


def apply(cfg):
    """Set the tuned configuration process-wide; returns the previous one."""
    old = {}
    for dt, (kr, ks) in _KEYS.items():
        if dt in cfg.get("panel", {}):
            p = cfg["panel"][dt]
            old.setdefault("panel", {})[dt] = {"rows": _lib.set_config(kr, int(p["rows"])),
                                               "strip": _lib.set_config(ks, int(p["strip"]))}
    if "batch_classes" in cfg:
        old["batch_classes"] = _lib.set_config(_lib.NLA_CFG_BATCH_CLASSES, int(cfg["batch_classes"]))
    return old


def _best(timings):
    """The candidate with the least time (ties: the first listed)."""
    return min(timings, key=lambda kv: kv[1])[0]


def tune_panels(dtype, measure=None, rows=PANEL_ROWS, strips=PANEL_STRIPS):
    """Time every (rows, strip) pair with measure(rows, strip) -> ms (default:
    TRSM + TRMM of the n = 8192 drop-in case on the GPU) and return
    {"rows", "strip", "ms": {"rows/strip": ms}} for the fastest."""
    kr, ks = _KEYS[dtype]
    if measure is None:
        measure = _panel_measure(dtype)
    old = (_lib.set_config(kr, rows[0]), _lib.set_config(ks, strips[0]))
    timings = []
    try:
        for r in rows:
            for s in strips:
                if _lib.set_config(kr, r) < 0 or _lib.set_config(ks, s) < 0:
                    continue
                timings.append(((r, s), float(measure(r, s))))
    finally:
        _lib.set_config(kr, old[0])
        _lib.set_config(ks, old[1])
    r, s = _best(timings)
    return {"rows": r, "strip": s, "ms": {f"{a}/{b}": round(t, 4) for (a, b), t in timings}}


def tune_batch(dtype, measure=None, masks=BATCH_MASKS):
    """Time every class mask with measure(mask) -> ms (default: a mixed batch
    of 2,000 systems, n log-uniform on [16, 1024], RHS on [1, 256])."""
    if measure is None:
        measure = _batch_measure(dtype)
    old = _lib.set_config(_lib.NLA_CFG_BATCH_CLASSES, masks[0])
    timings = []
    try:
        for mk in masks:
            _lib.set_config(_lib.NLA_CFG_BATCH_CLASSES, mk)
            timings.append((mk, float(measure(mk))))
    finally:
        _lib.set_config(_lib.NLA_CFG_BATCH_CLASSES, old)
    return {"mask": _best(timings), "ms": {hex(k): round(t, 4) for k, t in timings}}


def tune(device=None, force=False, dtypes=("f32", "f64"), save=True, apply_result=True):
    """Measure (or load from the cache) and, unless apply_result is False,
    apply the configuration for the current device; returns it."""
    import torch
    dev = torch.device(device or "cuda")
    props = torch.cuda.get_device_properties(dev)
    key = f"{props.name}|{getattr(props, 'gcnArchName', '')}|{props.multi_processor_count}"
    path = cache_path()
    cache = {}
    if os.path.exists(path):
        try:
            cache = json.load(open(path))
        except (OSError, ValueError):
            cache = {}
    cfg = cache.get(key)
    if cfg is None or force:
        cfg = {"panel": {}, "batch": {}}
        for dt in dtypes:
            cfg["panel"][dt] = tune_panels(dt)
        res = tune_batch("f64")
        cfg["batch"]["f64"] = res
        cfg["batch_classes"] = res["mask"]
        if save:
            cache[key] = cfg
            os.makedirs(os.path.dirname(path), exist_ok=True)
            tmp = path + ".tmp"
            with open(tmp, "w") as f:
                json.dump(cache, f, indent=1)
            os.replace(tmp, path)
    if apply_result:
        apply(cfg)
    return cfg


# ---- default measurements (GPU) -------------------------------------------

def _events_ms(torch, fn, reps):
    s = torch.cuda.current_stream()
    best = math.inf
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(s)
        fn()
        e1.record(s)
        torch.cuda.synchronize()
        best = min(best, e0.elapsed_time(e1))
    return best


def _panel_measure(dtype, n=8192, m=8192, reps=3):
    import torch
    from . import inputs, unified_rectrxm_
    dt = torch.float64 if dtype == "f64" else torch.float32
    A, B0 = inputs.device_workload1(n, m, "L", "L", dt, seed=1)
    B = B0.clone()

    def measure(rows, strip):
        total = 0.0
        for func in ("S", "M"):
            def one():
                unified_rectrxm_("L", "L", "N", 1.5, func, A, B)
            B.copy_(B0)
            one()
            torch.cuda.synchronize()
            total += _events_ms(torch, one, reps)
        return total
    return measure


def _batch_measure(dtype, count=2000, seed=0, reps=3):
    import numpy as np
    import torch
    from .batch import BatchPlan
    dt = torch.float64 if dtype == "f64" else torch.float32
    rng = np.random.default_rng(seed)
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    ns = np.exp(rng.uniform(np.log(16), np.log(1024), count)).astype(int)
    ms = np.exp(rng.uniform(np.log(1), np.log(256), count)).astype(int)
    As, Bs = [], []
    for n, m in zip(ns, ms):
        A = torch.empty((n, n), dtype=dt, device="cuda").t()
        A.uniform_(-1.0, 1.0, generator=g).mul_(1.0 / n)
        A.diagonal().uniform_(1.0, 2.0, generator=g)
        As.append(A)
        Bs.append(torch.empty((m, n), dtype=dt, device="cuda").t().normal_(generator=g))

    def measure(mask):
        plan = BatchPlan("L", "L", "N", "S", As, Bs, 1.0)  # reads the class mask now in force
        try:
            plan.execute()
            torch.cuda.synchronize()
            return _events_ms(torch, plan.execute, reps)
        finally:
            plan.close()
    return measure
