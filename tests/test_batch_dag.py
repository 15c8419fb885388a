"""Host-side replay of the persistent batched solver's task graph
(nla_batch_plan_describe): the data-race and ordering check the batched
kernels rely on, run on the CPU.

The plan's launches run in order; the tasks of one launch run concurrently,
in whatever order the persistent work-groups take them from the queue.  For
random batches of every side/uplo/trans x TRSM/TRMM (systems above 512 rows
solved in stages with grouped GEMMs, several 16-column strips and 64-column
GEMM tiles per system) the replay checks that
  * no two tasks of one launch write the same element of B, and no task
    reads an element another task of its launch writes;
  * every panel strip of a TRSM reads 64-block inverses made in an earlier
    launch;
  * executing the tasks with numpy, launch by launch, each launch's tasks in
    a random order, gives op(A)^-1 alpha B / alpha op(A) B for every system
    (against scipy's solve_triangular and numpy's product);
also for plans built with a perturbed task order (NLA_CFG_BATCH_ORDER), and
a planted mistake (a GEMM tile moved into the launch of the panel it
depends on, or a task dropped) is caught.
"""
import ctypes
import itertools

import numpy as np
import pytest
from scipy.linalg import solve_triangular

COMBOS = list(itertools.product("LR", "LU", "NT"))


def lib_and_types():
    from nextla_amd import _lib
    return _lib, _lib.load()


def canonical(side, uplo, trans, A):
    """op(A) for side 'L', op(A)^T for side 'R', and whether it is lower."""
    op = A.T if trans != "N" else A
    ue = uplo if trans == "N" else ("U" if uplo == "L" else "L")
    if side == "L":
        return op, ue == "L"
    return op.T, ue == "U"


def make_batch(rng, count, side, uplo, trans):
    systems = []
    for k in range(count):
        # random sizes, plus two systems that are solved in stages
        n = {0: 1300, 1: 700}.get(k, int(np.exp(rng.uniform(np.log(5), np.log(1500)))))
        m = {0: 70, 1: 33}.get(k, int(np.exp(rng.uniform(np.log(1), np.log(90)))))
        A = rng.uniform(-1, 1, (n, n)) / np.sqrt(n) + np.diag(2.0 + rng.random(n))
        A = np.tril(A) if uplo == "L" else np.triu(A)
        Bc = rng.standard_normal((n, m))  # canonical right-hand sides
        systems.append((n, m, A, Bc, float(rng.choice([1.0, -0.5, 1.5]))))
    return systems


def describe(side, uplo, trans, func, systems, dtype=1):
    _lib, lib = lib_and_types()
    items = (_lib.BatchItem * len(systems))()
    for i, (n, m, A, Bc, alpha) in enumerate(systems):
        items[i].n, items[i].m = n, m
        items[i].A, items[i].lda = 4096 * (i + 1), n  # addresses are never read
        items[i].B, items[i].ldb = 4096 * (i + 1) + 2048, n if side == "L" else m
        items[i].alpha = alpha
    total = lib.nla_batch_plan_describe(_lib.ch(side), _lib.ch(uplo), _lib.ch(trans), _lib.ch(func), dtype, items,
                                        len(systems), None, 0)
    assert total >= 0, total
    out = (_lib.BatchTaskDesc * max(total, 1))()
    got = lib.nla_batch_plan_describe(_lib.ch(side), _lib.ch(uplo), _lib.ch(trans), _lib.ch(func), dtype, items,
                                      len(systems), out, total)
    assert got == total
    return [dict(launch=d.launch, kind=d.kind, system=d.system, r0=d.r0, rows=d.rows, c0=d.c0, cols=d.cols,
                 k0=d.k0, kn=d.kn, alpha=d.alpha, beta=d.beta) for d in out[:total]]


def footprint(t):
    """(writes, reads) as lists of (system, r0, r1, c0, c1) boxes of B; the
    inversion writes its block of inverses (negative row marker)."""
    s = t["system"]
    if t["kind"] == 0:
        return [(s, -1 - t["r0"], -t["r0"], 0, 1)], []
    box = (s, t["r0"], t["r0"] + t["rows"], t["c0"], t["c0"] + t["cols"])
    if t["kind"] == 1:
        return [box], [box]
    reads = [(s, t["k0"], t["k0"] + t["kn"], t["c0"], t["c0"] + t["cols"])]
    if t["beta"] != 0.0:
        reads.append(box)
    return [box], reads


def overlap(a, b):
    return a[0] == b[0] and a[1] < b[2] and b[1] < a[2] and a[3] < b[4] and b[3] < a[4]


def check_races(tasks):
    """Conflicts between tasks of one launch (pairs within one system)."""
    bad = []
    by_launch = {}
    for t in tasks:
        by_launch.setdefault(t["launch"], {}).setdefault(t["system"], []).append(t)
    for launch, per_sys in by_launch.items():
        for ts in per_sys.values():
            fps = [footprint(t) for t in ts]
            for a in range(len(ts)):
                for b in range(len(ts)):
                    if a == b:
                        continue
                    wa, ra = fps[a]
                    wb, _ = fps[b]
                    if any(overlap(x, y) for x in wa for y in wb) or any(overlap(x, y) for x in ra for y in wb):
                        bad.append((launch, ts[a], ts[b]))
    return bad


def replay(tasks, systems, lower_of, func, rng):
    """Execute the tasks with numpy in launch order, each launch shuffled."""
    B = [Bc.copy() for (_, _, _, Bc, _) in systems]
    Mc = [lower_of[i][0] for i in range(len(systems))]
    lower = [lower_of[i][1] for i in range(len(systems))]
    inv_ready = set()
    for launch in sorted({t["launch"] for t in tasks}):
        ts = [t for t in tasks if t["launch"] == launch]
        rng.shuffle(ts)
        made = []
        for t in ts:
            s = t["system"]
            R = slice(t["r0"], t["r0"] + t["rows"])
            C = slice(t["c0"], t["c0"] + t["cols"])
            if t["kind"] == 0:
                made.append((s, t["r0"]))
            elif t["kind"] == 1:
                blk = Mc[s][R, R]
                blk = np.tril(blk) if lower[s] else np.triu(blk)
                if func == "S":
                    for r in range(t["r0"], t["r0"] + t["rows"], 64):
                        assert (s, r) in inv_ready, ("panel before its inverse", t)
                    B[s][R, C] = t["alpha"] * solve_triangular(blk, B[s][R, C], lower=lower[s])
                else:
                    B[s][R, C] = t["alpha"] * (blk @ B[s][R, C])
            else:
                K = slice(t["k0"], t["k0"] + t["kn"])
                B[s][R, C] = t["beta"] * B[s][R, C] + t["alpha"] * (Mc[s][R, K] @ B[s][K, C])
        inv_ready.update(made)  # visible from the next launch on
    return B


def expected(systems, lower_of, func):
    out = []
    for i, (n, m, A, Bc, alpha) in enumerate(systems):
        M, low = lower_of[i]
        T = np.tril(M) if low else np.triu(M)
        out.append(alpha * (solve_triangular(T, Bc, lower=low) if func == "S" else T @ Bc))
    return out


def run_case(side, uplo, trans, func, seed, order=0):
    _lib, lib = lib_and_types()
    rng = np.random.default_rng(seed)
    systems = make_batch(rng, 14, side, uplo, trans)
    old = _lib.set_config(_lib.NLA_CFG_BATCH_ORDER, order)
    try:
        tasks = describe(side, uplo, trans, func, systems)
    finally:
        _lib.set_config(_lib.NLA_CFG_BATCH_ORDER, old)
    lower_of = [canonical(side, uplo, trans, A) for (_, _, A, _, _) in systems]
    return rng, systems, lower_of, tasks


@pytest.mark.parametrize("side,uplo,trans", COMBOS)
@pytest.mark.parametrize("func", ["S", "M"])
def test_task_graph_race_free_and_correct(nla, side, uplo, trans, func):
    import zlib
    rng, systems, lower_of, tasks = run_case(side, uplo, trans, func, seed=zlib.crc32(f"{side}{uplo}{trans}{func}".encode()))
    assert any(t["kind"] == 2 for t in tasks), "no staged system in the batch"
    assert not check_races(tasks)
    want = expected(systems, lower_of, func)
    for rep in range(2):
        got = replay(tasks, systems, lower_of, func, np.random.default_rng(rep))
        for g, w in zip(got, want):
            assert np.linalg.norm(g - w) <= 1e-12 * max(np.linalg.norm(w), 1.0)
    # every element of every system is written by exactly the tasks the
    # recursion would: each panel row block is solved once per strip
    panels = {}
    for t in tasks:
        if t["kind"] == 1:
            key = (t["system"], t["r0"], t["c0"])
            panels[key] = panels.get(key, 0) + 1
    assert all(v == 1 for v in panels.values())


@pytest.mark.parametrize("side,uplo,trans,func", [("L", "L", "N", "S"), ("R", "U", "T", "M"), ("L", "U", "T", "S")])
def test_perturbed_task_order(nla, side, uplo, trans, func):
    """A plan built with a pseudo-random task order within each launch is
    still race-free and correct (the device kernels take tasks in this
    order; tests/test_gpu_batch.py checks their bits do not change)."""
    _, systems, lower_of, t0 = run_case(side, uplo, trans, func, seed=5, order=0)
    _, _, _, t1 = run_case(side, uplo, trans, func, seed=5, order=7)
    key = lambda t: (t["launch"], t["kind"], t["system"], t["r0"], t["c0"])
    assert sorted(map(key, t0)) == sorted(map(key, t1)) and list(map(key, t0)) != list(map(key, t1))
    assert not check_races(t1)
    want = expected(systems, lower_of, func)
    got = replay(t1, systems, lower_of, func, np.random.default_rng(3))
    for g, w in zip(got, want):
        assert np.linalg.norm(g - w) <= 1e-12 * max(np.linalg.norm(w), 1.0)


def test_planted_mistakes_are_caught(nla):
    """A GEMM tile moved into the launch of the panel it reads races with it;
    a dropped task leaves a wrong result."""
    _, systems, lower_of, tasks = run_case("L", "L", "N", "S", seed=11)
    g = next(t for t in tasks if t["kind"] == 2)
    moved = [dict(t) for t in tasks]
    i = tasks.index(g)
    # into the launch of the panel whose rows it reads
    moved[i]["launch"] = next(t["launch"] for t in tasks
                              if t["kind"] == 1 and t["system"] == g["system"] and t["r0"] == g["k0"])
    assert check_races(moved)
    dropped = [t for t in tasks if t is not g]
    want = expected(systems, lower_of, "S")
    got = replay(dropped, systems, lower_of, "S", np.random.default_rng(0))
    assert max(np.linalg.norm(a - b) / np.linalg.norm(b) for a, b in zip(got, want)) > 1e-6
