"""Differentiable 3D NLM: a torch.autograd.Function over the C-ABI forward
(gnlm3d_forward) and the hand-written adjoint (gnlm3d_backward).

The forward keeps y and Z (two floats per voxel) for the backward; the adjoint
kernel recomputes the patch distances instead of storing the voxels x offsets
weight tensor.  Gradients flow to the noisy input u and to h.

h enters the C-ABI by value (the kernels fold -log2(e)/(P h^2) on the host), so
its value must be on the host when the forward is enqueued:
  * a Python float, or a 0-d CPU tensor, costs no device synchronisation;
    the gradient of a CPU tensor h is copied back to the host in the backward
    (one 4-byte device-to-host copy, which waits for the adjoint kernels);
  * a 0-d CUDA tensor works too, but reading its value waits for the device
    on every forward.
The gradient dL/dh is returned with h's shape and on h's device.
"""
import torch

from . import _check_dev, _check_like, _check_vol, nlm3d as _fwd, nlm3d_backward as _bwd


class NLM3DFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u, h, r_p, r_s):
        _check_vol(u, 3)
        _check_dev([u], ["u"])
        u = u.contiguous()
        hv = float(h)
        y, Z = _fwd(u, r_p, r_s, hv, want_Z=True)
        ctx.save_for_backward(u, y, Z)
        ctx.hv, ctx.r_p, ctx.r_s = hv, int(r_p), int(r_s)
        ctx.h_like = h if isinstance(h, torch.Tensor) else None
        return y

    @staticmethod
    def backward(ctx, gy):
        u, y, Z = ctx.saved_tensors
        gy = gy.contiguous()
        _check_like(gy, u, "grad_output")
        gu, gh = _bwd(u, y, Z, gy, ctx.r_p, ctx.r_s, ctx.hv)
        gh_out = None
        if ctx.h_like is not None and ctx.needs_input_grad[1]:
            gh_out = gh.reshape(ctx.h_like.shape).to(ctx.h_like.device)
        return (gu if ctx.needs_input_grad[0] else None), gh_out, None, None


def nlm3d_autograd(u, h, r_p, r_s):
    """y = NLM(u; h, r_p, r_s), differentiable in u and h (see the module notes
    on where h should live)."""
    return NLM3DFunction.apply(u, h, r_p, r_s)
