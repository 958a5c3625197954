"""torch.autograd.gradcheck of the implicit momentum residual through the
C-ABI JVP/VJP, on a 5x5x5-cell C2-type grid (immersed + LES flags on).

Run as its own process (the library is loaded before torch):
    python tests/gradcheck_snes.py
Prints "gradcheck ok" and exits 0 on success.
"""
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "vfs-geophysics-3.1_amd"))

import numpy as np  # noqa: E402

import vfsgpu as V  # noqa: E402
from vfsgpu import workloads as W  # noqa: E402


def main():
    V.lib()
    import torch
    from vfsgpu.autograd import SnesProblem, residual_function

    cfg, spec = W.make_config("C2", (5, 5, 5), dt=0.05)
    ctx = V.Ctx(cfg)
    ctx.build_metrics(V.DeviceArray.from_host(W.grid(spec)))
    # a mask with a solid cell, its IB ring and fluid, so every face rule runs
    nv = np.zeros((cfg["mz"], cfg["my"], cfg["mx"]))
    nv[1:6, 1:6, 1:6] = 0.0
    nv[2:5, 2:5, 2:5] = 1.0
    nv[3, 3, 3] = 3.0
    ctx.set_nvert(V.DeviceArray.from_host(nv.ravel()))
    P = ctx.P
    U0 = W.random_vector(cfg, 0.05, 51)
    ucat = W.random_vector(cfg, 0.5, 52)
    nut = np.abs(W.random_vector(cfg, 1e-3, 53)[:P])
    prob = SnesProblem(ctx, W.random_vector(cfg, 0.05, 54), W.random_vector(cfg, 0.01, 55),
                       W.random_vector(cfg, 0.01, 56), nut, ucat)
    F = residual_function()
    U = torch.from_numpy(U0.copy()).requires_grad_(True)
    # the VJP is deterministic (a gather with a fixed summation order), so
    # gradcheck's reentrancy check runs with its default nondet_tol of 0
    ok = torch.autograd.gradcheck(lambda u: F.apply(u, prob), (U,), eps=1e-6, atol=1e-5, rtol=1e-4,
                                  check_forward_ad=True, check_backward_ad=True, fast_mode=False)
    print("gradcheck ok" if ok else "gradcheck FAILED", flush=True)
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
