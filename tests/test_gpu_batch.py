"""Batched solves / multiplies (nla_batch_*): parity with one plain call per
system, the oracle's error bounds, bitwise equality of a system alone and
inside a mixed batch, and hipGraph replays against the eager launch."""
import itertools

import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu
DEV = "cuda"


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a).T)).to(DEV).t()


def make_system(rng, n, m, side, uplo, dt):
    A = np.tril(rng.uniform(-1, 1, (n, n)) / max(n, 1)) + np.diag(1 + rng.random(n))
    if uplo == "U":
        A = A.T.copy()
    B = rng.standard_normal((n, m) if side == "L" else (m, n))
    return A.astype(dt), B.astype(dt)


SHAPES = [(1, 1), (16, 5), (64, 16), (65, 33), (100, 1), (200, 100), (512, 17), (700, 40), (1024, 3)]


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_batch_all_combos(nla, dt):
    from nextla_amd import batch
    rng = np.random.default_rng(5)
    eps = np.finfo(dt).eps
    for side, uplo, trans, func in itertools.product("LR", "LU", "NT", "SM"):
        hostA, hostB, As, Bs, refs = [], [], [], [], []
        for n, m in SHAPES:
            A, B = make_system(rng, n, m, side, uplo, dt)
            hostA.append(A)
            hostB.append(B)
            As.append(dev(A))
            Bs.append(dev(B))
            # one plain call per system (the reference's usage pattern)
            Bp = dev(B)
            nla.unified_rectrxm_(side, uplo, trans, 1.25, func, dev(A), Bp)
            refs.append(Bp)
        batch.batched_rectrxm_(side, uplo, trans, func, As, Bs, alphas=1.25)
        torch.cuda.synchronize()
        for A, B, X, R in zip(hostA, hostB, Bs, refs):
            n = A.shape[0]
            Xh = X.cpu().numpy()
            e = (O.backward_error if func == "S" else O.product_error)(side, uplo, trans, 1.25, A, B, Xh)
            assert e <= 10 * n * eps, (side, uplo, trans, func, n, e)
            assert O.rel_error(Xh, R.cpu().numpy()) <= 100 * n * eps, (side, uplo, trans, func, n)


def test_alone_equals_in_mix(nla):
    """A system's result does not depend on what else is in the batch."""
    from nextla_amd import batch
    rng = np.random.default_rng(6)
    A, B = make_system(rng, 300, 45, "L", "L", np.float64)
    alone = [dev(B)]
    batch.batched_rectrxm_("L", "L", "N", "S", [dev(A)], alone)
    As, Bs = [], []
    for n, m in [(1000, 7), (300, 45), (33, 200), (512, 512), (5, 5)]:
        if (n, m) == (300, 45):
            As.append(dev(A))
            Bs.append(dev(B))
            mine = len(Bs) - 1
        else:
            a, b = make_system(rng, n, m, "L", "L", np.float64)
            As.append(dev(a))
            Bs.append(dev(b))
    batch.batched_rectrxm_("L", "L", "N", "S", As, Bs)
    torch.cuda.synchronize()
    assert torch.equal(alone[0], Bs[mine])


def test_graph_replays_match_eager(nla):
    from nextla_amd import batch
    rng = np.random.default_rng(7)
    systems = [make_system(rng, n, m, "L", "L", np.float64) for n, m in SHAPES]
    As = [dev(a) for a, _ in systems]
    B0 = [dev(b) for _, b in systems]
    Bs = [b.clone() for b in B0]
    plan = batch.BatchPlan("L", "L", "N", "S", As, Bs, 0.5)
    plan.execute(use_graph=False)
    torch.cuda.synchronize()
    eager = [b.clone() for b in Bs]
    for rep in range(4):
        for b, b0 in zip(Bs, B0):
            b.copy_(b0)  # the program's own eager work between replays
        torch.cuda.synchronize()
        plan.execute(use_graph=True)
        torch.cuda.synchronize()
        for b, e in zip(Bs, eager):
            assert torch.equal(b, e), rep
    plan.close()


BIG = [(1025, 3), (1100, 40), (1600, 17), (2100, 70)]


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_staged_systems_all_combos(nla, dt):
    """Systems above 1024 rows are solved in 512-row stages (panel launch +
    grouped GEMM per stage), mixed with single-panel systems in one plan:
    every combo within the oracle's bounds and agreeing with plain calls."""
    from nextla_amd import batch
    rng = np.random.default_rng(23)
    eps = np.finfo(dt).eps
    shapes = BIG + [(64, 16), (700, 9)]
    for side, uplo, trans, func in itertools.product("LR", "LU", "NT", "SM"):
        hostA, hostB, As, Bs, refs = [], [], [], [], []
        for n, m in shapes:
            A, B = make_system(rng, n, m, side, uplo, dt)
            hostA.append(A)
            hostB.append(B)
            As.append(dev(A))
            Bs.append(dev(B))
            Bp = dev(B)
            nla.unified_rectrxm_(side, uplo, trans, -0.75, func, dev(A), Bp)
            refs.append(Bp)
        batch.batched_rectrxm_(side, uplo, trans, func, As, Bs, alphas=-0.75)
        torch.cuda.synchronize()
        for A, B, X, R in zip(hostA, hostB, Bs, refs):
            n = A.shape[0]
            Xh = X.cpu().numpy()
            e = (O.backward_error if func == "S" else O.product_error)(side, uplo, trans, -0.75, A, B, Xh)
            assert e <= 10 * n * eps, (side, uplo, trans, func, n, e)
            assert O.rel_error(Xh, R.cpu().numpy()) <= 100 * n * eps, (side, uplo, trans, func, n)


def test_staged_alone_equals_in_mix_and_graph(nla):
    """A staged system gets the same bits alone, inside a mixed batch, and
    from graph replays of the plan."""
    from nextla_amd import batch
    rng = np.random.default_rng(29)
    for side, uplo, func in (("L", "L", "S"), ("R", "U", "S"), ("L", "U", "M"), ("R", "L", "M")):
        A, B = make_system(rng, 1600, 33, side, uplo, np.float64)
        alone = [dev(B)]
        batch.batched_rectrxm_(side, uplo, "N", func, [dev(A)], alone, alphas=1.5)
        others = [make_system(rng, n, m, side, uplo, np.float64) for n, m in ((2500, 5), (90, 20), (1030, 64))]
        As = [dev(a) for a, _ in others] + [dev(A)]
        B0 = [dev(b) for _, b in others] + [dev(B)]
        Bs = [b.clone() for b in B0]
        plan = batch.BatchPlan(side, uplo, "N", func, As, Bs, 1.5)
        plan.execute()
        torch.cuda.synchronize()
        eager = [b.clone() for b in Bs]
        assert torch.equal(eager[-1], alone[0]), (side, uplo, func)
        for rep in range(3):
            for b, b0 in zip(Bs, B0):
                b.copy_(b0)
            plan.execute(use_graph=True)
            torch.cuda.synchronize()
            for b, e in zip(Bs, eager):
                assert torch.equal(b, e), (side, uplo, func, rep)
        plan.close()


@pytest.mark.parametrize("mask", [0x8, 0xA, 0xF, 0x18, 0x1A])
def test_size_classes_bitwise_identical(nla, mask):
    """Whatever size classes a plan may use (NLA_CFG_BATCH_CLASSES), every
    system gets the same bits: the classes differ only in LDS stride (with
    and without bit 0x10, which changes the algorithm above 512 rows)."""
    from nextla_amd import _lib, batch
    shapes = [(16, 3), (128, 20), (129, 7), (256, 33), (300, 16), (512, 5), (513, 40), (1024, 17)]
    results = {}
    for cur in (0xF | (mask & 0x10), mask):
        old = _lib.set_config(_lib.NLA_CFG_BATCH_CLASSES, cur)
        try:
            for side, uplo, func in (("L", "L", "S"), ("R", "U", "M")):
                rng2 = np.random.default_rng(3)
                As, Bs = [], []
                for n, m in shapes:
                    A, B = make_system(rng2, n, m, side, uplo, np.float64)
                    As.append(dev(A))
                    Bs.append(dev(B))
                batch.batched_rectrxm_(side, uplo, "N", func, As, Bs, alphas=0.75)
                torch.cuda.synchronize()
                results.setdefault((side, uplo, func), []).append([b.cpu().numpy() for b in Bs])
        finally:
            _lib.set_config(_lib.NLA_CFG_BATCH_CLASSES, old)
    for key, (ref, got) in results.items():
        for r, g in zip(ref, got):
            assert np.array_equal(r, g), key


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_staged_above_512_option(nla, dt):
    """Class-mask bit 0x10: systems of 513..1024 rows solved in two 512-row
    stages (panel, grouped GEMM, panel) instead of one 1024-row panel.  Every
    combo stays within the oracle's bounds, and a system alone and in a
    mixed batch gets the same bits under that option."""
    from nextla_amd import _lib, batch
    rng = np.random.default_rng(17)
    eps = np.finfo(dt).eps
    shapes = [(513, 40), (700, 9), (1024, 17), (300, 5), (64, 1)]
    old = _lib.set_config(_lib.NLA_CFG_BATCH_CLASSES, 0x1F)
    try:
        for side, uplo, trans, func in itertools.product("LR", "LU", "NT", "SM"):
            hostA, hostB, As, Bs = [], [], [], []
            for n, m in shapes:
                A, B = make_system(rng, n, m, side, uplo, dt)
                hostA.append(A)
                hostB.append(B)
                As.append(dev(A))
                Bs.append(dev(B))
            batch.batched_rectrxm_(side, uplo, trans, func, As, Bs, alphas=-0.75)
            torch.cuda.synchronize()
            for A, B, X in zip(hostA, hostB, Bs):
                n = A.shape[0]
                e = (O.backward_error if func == "S" else O.product_error)(side, uplo, trans, -0.75, A, B,
                                                                          X.cpu().numpy())
                assert e <= 10 * n * eps, (side, uplo, trans, func, n, e)
            if (side, uplo, trans) == ("L", "L", "N"):
                alone = [dev(hostB[2])]
                batch.batched_rectrxm_(side, uplo, trans, func, [As[2]], alone, alphas=-0.75)
                torch.cuda.synchronize()
                assert torch.equal(alone[0], Bs[2]), func
    finally:
        _lib.set_config(_lib.NLA_CFG_BATCH_CLASSES, old)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("side,uplo,trans,func", [("L", "L", "N", "S"), ("R", "U", "T", "M"), ("L", "U", "T", "S"),
                                                  ("R", "L", "N", "S")])
def test_perturbed_task_order_bitwise(nla, dt, side, uplo, trans, func):
    """Plans whose tasks run in pseudo-random orders within each launch
    (NLA_CFG_BATCH_ORDER; the host replay in tests/test_batch_dag.py checks
    the same plans for races) give the same bits as the default order: no
    task depends on which work-group runs it or when."""
    from nextla_amd import _lib, batch
    shapes = [(16, 3), (128, 20), (300, 40), (513, 40), (700, 70), (1300, 33), (64, 100)]
    out = []
    for order in (0, 3, 11):
        rng = np.random.default_rng(17)
        As, Bs = [], []
        for n, m in shapes:
            A, B = make_system(rng, n, m, side, uplo, dt)
            As.append(dev(A))
            Bs.append(dev(B))
        old = _lib.set_config(_lib.NLA_CFG_BATCH_ORDER, order)
        try:
            batch.batched_rectrxm_(side, uplo, trans, func, As, Bs, alphas=-0.5)
        finally:
            _lib.set_config(_lib.NLA_CFG_BATCH_ORDER, old)
        torch.cuda.synchronize()
        out.append([b.cpu().numpy() for b in Bs])
    for got in out[1:]:
        for r, g in zip(out[0], got):
            assert np.array_equal(r, g)


def test_launch_blocking_bitwise(nla, tmp_path):
    """The same work with every launch serialized (HIP_LAUNCH_BLOCKING=1, in
    a child process: the variable is read when HIP starts) gives the same
    bits: the plain recursion, a batch, the refined solve and the int8
    residual do not depend on launches overlapping."""
    import os
    import subprocess
    import sys
    here = os.path.dirname(os.path.abspath(__file__))
    child = os.path.join(here, "launch_blocking_child.py")
    env = dict(os.environ, HIP_LAUNCH_BLOCKING="1")
    blocked = tmp_path / "blocked.npz"
    r = subprocess.run([sys.executable, child, str(blocked)], env=env, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    here_out = tmp_path / "async.npz"
    import importlib.util
    spec = importlib.util.spec_from_file_location("launch_blocking_child", child)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    mod.run(str(here_out))
    a, b = np.load(blocked), np.load(here_out)
    assert sorted(a.files) == sorted(b.files) and len(a.files) >= 4
    for k in a.files:
        assert np.array_equal(a[k], b[k]), k
