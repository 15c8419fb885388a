"""A fixed set of library calls whose outputs tests/test_gpu_batch.py
compares between a process with HIP_LAUNCH_BLOCKING=1 (every launch
serialized) and one without.  Run as a script: python launch_blocking_child.py
OUT.npz."""
import os
import sys

import numpy as np


def run(out_path):
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    for p in (root, os.path.join(root, "nextla.jl_amd")):
        if p not in sys.path:
            sys.path.insert(0, p)
    import torch
    import nextla_amd as nla
    from nextla_amd import batch
    from nextla_amd.mixed import rectrxm_mixed_, residual
    nla.load()
    g = torch.Generator(device="cuda")
    g.manual_seed(123)

    def tri(n):
        a = torch.empty((n, n), dtype=torch.float64, device="cuda").t()
        a.uniform_(-1.0, 1.0, generator=g).mul_(n ** -0.5)
        a.copy_(torch.tril(a))
        a.diagonal().add_(2.0)
        return a

    def rhs(n, m):
        return torch.empty((m, n), dtype=torch.float64, device="cuda").t().normal_(generator=g)

    out = {}
    A = tri(2304)
    B = rhs(2304, 300)
    for func in "SM":
        X = B.clone()
        nla.unified_rectrxm_("L", "L", "N", 1.5, func, A, X)
        out["rectrxm_" + func] = X
    As = [tri(n) for n in (40, 300, 700, 1300)]
    Bs = [rhs(a.shape[0], m) for a, m in zip(As, (3, 40, 70, 17))]
    batch.batched_rectrxm_("L", "L", "N", "S", As, Bs, alphas=0.75)
    for i, b in enumerate(Bs):
        out[f"batch_{i}"] = b
    X = B.clone()
    rectrxm_mixed_("L", "L", "N", 1.0, A, X, precision="auto", tol=1e-14)
    out["mixed"] = X
    Xr = B.clone()
    R = torch.empty_like(B)
    residual("L", "L", "N", 1.0, A, Xr, B, R, method="int8")
    out["residual"] = R
    torch.cuda.synchronize()
    np.savez(out_path, **{k: v.cpu().numpy() for k, v in out.items()})


if __name__ == "__main__":
    run(sys.argv[1])
