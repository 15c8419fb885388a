"""Writes tests/golden/blas_fixtures.npz: seeded inputs and the outputs of the
reference tests' oracle, LinearAlgebra.BLAS.trsm!/trmm! (test/unified_rectrxm.jl:36-40),
i.e. OpenBLAS 0.3.29 (test/Manifest.toml:494) — here the same library build
through scipy.linalg.blas.

The reference itself is Julia and cannot run in this environment; its test
suite holds no golden vectors (inputs there are unseeded `rand`), so these
fixtures pin the library oracle those tests compare against.

Run:  python tests/golden/make_golden.py
"""
import itertools
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "nextla.jl_amd"))

from nextla_amd import inputs  # noqa: E402
from scipy.linalg import blas as sb  # noqa: E402

COMBOS = list(itertools.product("LR", "LU", "NTC", "SM"))


def blas(side, uplo, trans, alpha, func, A, B):
    pre = "d" if A.dtype == np.float64 else "s"
    fn = getattr(sb, pre + ("trsm" if func == "S" else "trmm"))
    return fn(alpha, A, B, side=0 if side == "L" else 1, lower=1 if uplo == "L" else 0,
              trans_a={"N": 0, "T": 1, "C": 2}[trans], diag=0)


def main():
    out = {}
    manifest = []
    rng = np.random.Generator(np.random.PCG64(20250101))

    # 1. The reference test grid (test/unified_rectrxm.jl:11-43) at n in {16, 32},
    #    m in {1, 8}: FP64, alpha = 1, A = tri(U(1,2)) + 10I, B = U(1,2).
    for n in (16, 32):
        for uplo in "LU":
            out[f"grid/A/{n}/{uplo}"] = inputs.reference_test_tri(n, uplo, rng)
        for m in (1, 8):
            for side in "LR":
                shape = (n, m) if side == "L" else (m, n)
                out[f"grid/B/{n}/{m}/{side}"] = inputs.reference_test_rhs(*shape, rng)
            for side, uplo, trans, func in COMBOS:
                A = out[f"grid/A/{n}/{uplo}"]
                B = out[f"grid/B/{n}/{m}/{side}"]
                key = f"grid/X/{n}/{m}/{side}{uplo}{trans}{func}"
                out[key] = blas(side, uplo, trans, 1.0, func, A, B)
                manifest.append({"key": key, "A": f"grid/A/{n}/{uplo}", "B": f"grid/B/{n}/{m}/{side}",
                                 "side": side, "uplo": uplo, "trans": trans, "func": func, "alpha": 1.0,
                                 "dtype": "float64"})

    # 2. alpha != 1 (never exercised by the reference tests) in both precisions,
    #    n = 40 (a non-power of two), m = 7, workload-style inputs.
    for dt, alpha in ((np.float64, 1.5), (np.float32, -0.75)):
        tag = "d" if dt == np.float64 else "s"
        n, m = 40, 7
        for uplo in "LU":
            out[f"alpha/A/{tag}/{uplo}"] = inputs.workload1_tri(n, uplo, rng, dt)
        for side in "LR":
            shape = (n, m) if side == "L" else (m, n)
            out[f"alpha/B/{tag}/{side}"] = inputs.normal_rhs(*shape, rng, dt)
        for side, uplo, trans, func in COMBOS:
            A = out[f"alpha/A/{tag}/{uplo}"]
            B = out[f"alpha/B/{tag}/{side}"]
            key = f"alpha/X/{tag}/{side}{uplo}{trans}{func}"
            out[key] = blas(side, uplo, trans, alpha, func, A, B)
            manifest.append({"key": key, "A": f"alpha/A/{tag}/{uplo}", "B": f"alpha/B/{tag}/{side}",
                             "side": side, "uplo": uplo, "trans": trans, "func": func, "alpha": alpha,
                             "dtype": np.dtype(dt).name})

    # 3. The 512 x 512 CPU-runnable workload: FP64 left-lower non-transposed
    #    TRSM and TRMM, alpha = 1, A off-diagonal U(-1,1)/512, diagonal 1+U(0,1)
    #    (seed 0), B 512 x 64 N(0,1).
    A = inputs.workload0_tri(512, seed=0)
    B = inputs.normal_rhs(512, 64, np.random.Generator(np.random.PCG64(1)))
    out["w0/A"] = A
    out["w0/B"] = B
    for func in "SM":
        key = f"w0/X/{func}"
        out[key] = blas("L", "L", "N", 1.0, func, A, B)
        manifest.append({"key": key, "A": "w0/A", "B": "w0/B", "side": "L", "uplo": "L", "trans": "N",
                         "func": func, "alpha": 1.0, "dtype": "float64"})

    np.savez_compressed(os.path.join(HERE, "blas_fixtures.npz"), **out)
    with open(os.path.join(HERE, "blas_fixtures.json"), "w") as f:
        json.dump({"library": "OpenBLAS 0.3.29 via scipy.linalg.blas", "cases": manifest}, f, indent=0)
    print(f"wrote {len(manifest)} cases")


if __name__ == "__main__":
    main()
