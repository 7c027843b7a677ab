"""torch.autograd.Function over the C ABI's forward / adjoint (checking only).

``VtiModeling.apply(vp, grad_ctx)`` maps a vp model tensor (float64, CPU)
to the shot record through ``vti_grad_forward``; its backward calls the
hand-written adjoint (``vti_grad_run`` with the incoming gradient as the
residual).  Used by the tests to run torch.autograd.gradcheck against the
HIP adjoint; the production gradient path is ``vti.Gradient.run``.
"""
import numpy as np
import torch

from . import vti


class VtiModeling(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vp, grad_obj, eps, dl, nsnap=0):
        ctx.grad_obj = grad_obj
        ctx.eps, ctx.dl, ctx.nsnap = eps, dl, nsnap
        ctx.save_for_backward(vp)
        grad_obj.set_model(vp.detach().cpu().numpy(), eps, dl, keep_tables=True)
        d = grad_obj.forward()
        return torch.from_numpy(d).to(vp.dtype)

    @staticmethod
    def backward(ctx, grad_out):
        (vp,) = ctx.saved_tensors
        g = ctx.grad_obj
        g.set_model(vp.detach().cpu().numpy(), ctx.eps, ctx.dl, keep_tables=True)
        r = grad_out.detach().cpu().numpy().astype(np.float64)
        out = g.run(resid=r, nsnap=ctx.nsnap)
        return torch.from_numpy(out["grad_vp"]).to(vp.dtype), None, None, None, None
