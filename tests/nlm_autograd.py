"""Test-side torch.autograd.Function over the C-ABI forward and adjoint.

The forward keeps y and Z (two floats per voxel) for the backward, which
recomputes the patch distances in the hand-written adjoint kernel.
"""
import torch

import gnlm


class NLM3D(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u, h, r_p, r_s):
        u = u.contiguous()
        hv = float(h)
        y, Z = gnlm.nlm3d(u, r_p, r_s, hv, want_Z=True)
        ctx.save_for_backward(u, y, Z)
        ctx.hv, ctx.r_p, ctx.r_s = hv, r_p, r_s
        ctx.h_shape = h.shape if isinstance(h, torch.Tensor) else None
        return y

    @staticmethod
    def backward(ctx, gy):
        u, y, Z = ctx.saved_tensors
        gu, gh = gnlm.nlm3d_backward(u, y, Z, gy.contiguous(), ctx.r_p, ctx.r_s, ctx.hv)
        gh_out = gh.reshape(ctx.h_shape) if ctx.h_shape is not None else None
        return gu, gh_out, None, None


def nlm3d(u, h, r_p, r_s):
    return NLM3D.apply(u, h, r_p, r_s)
