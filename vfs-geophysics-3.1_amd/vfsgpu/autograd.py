"""torch.autograd view of the implicit momentum residual F(U) of
FormFunction_SNES (momentum.c:2197-2301), for torch.autograd.gradcheck.

The tensors are CPU float64; every call copies U to the device and runs the
library (vfs_formfunction_snes, vfs_rhs_vjp for backward, vfs_rhs_jvp for
forward-mode AD).  torch is used only for the autograd bookkeeping here:
the product path is the C-ABI.

Import order: load the library (vfsgpu.lib()) before importing torch, so
that the process holds one HIP runtime (the library's).
"""
import numpy as np

import vfsgpu as V


class SnesProblem:
    """Fixed data of F: U_o, rhs_o, dP, nu_t and the frozen ucat."""

    def __init__(self, ctx, U_o, rhs_o, dP, nu_t, ucat):
        self.ctx = ctx
        self.U_o = V.DeviceArray.from_host(U_o)
        self.rhs_o = V.DeviceArray.from_host(rhs_o)
        self.dP = V.DeviceArray.from_host(dP)
        self.nu_t = None if nu_t is None else V.DeviceArray.from_host(nu_t)
        self.ucat0 = np.array(ucat, dtype=np.float64)
        self.dU = ctx.vec()
        self.dw = ctx.vec()
        self.out = ctx.vec()
        self.lu = ctx.vec()
        self.uc = ctx.vec()

    def residual(self, U):
        self.dU.upload(U)
        self.uc.upload(self.ucat0)
        self.ctx.formfunction_snes(self.dU, self.U_o, self.rhs_o, self.dP, self.nu_t, self.lu, self.uc, self.out)
        return self.out.host()

    def jvp(self, U, w):
        self.dU.upload(U)
        self.uc.upload(self.ucat0)
        self.dw.upload(w)
        self.ctx.rhs_jvp(self.dU, self.uc, self.nu_t, self.dw, self.out)
        return self.out.host()

    def vjp(self, U, v):
        self.dU.upload(U)
        self.uc.upload(self.ucat0)
        self.dw.upload(v)
        self.ctx.rhs_vjp(self.dU, self.uc, self.nu_t, self.dw, self.out)
        return self.out.host()


def residual_function():
    """Return the torch.autograd.Function class (imports torch lazily)."""
    import torch

    class SnesResidual(torch.autograd.Function):
        @staticmethod
        def forward(U, prob):
            return torch.from_numpy(prob.residual(U.detach().numpy()))

        @staticmethod
        def setup_context(ctx, inputs, output):
            U, prob = inputs
            ctx.prob = prob
            ctx.save_for_backward(U)
            ctx.save_for_forward(U)

        @staticmethod
        def backward(ctx, g):
            (U,) = ctx.saved_tensors
            jtv = ctx.prob.vjp(U.detach().numpy(), np.ascontiguousarray(g.detach().numpy()))
            return torch.from_numpy(jtv), None

        @staticmethod
        def jvp(ctx, w, _):
            (U,) = ctx.saved_tensors
            return torch.from_numpy(ctx.prob.jvp(U.detach().numpy(), np.ascontiguousarray(w.detach().numpy())))

    return SnesResidual
