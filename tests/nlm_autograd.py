"""The tests' handle on the package's differentiable 3D NLM
(gnlm.autograd.NLM3DFunction over the C-ABI forward and adjoint)."""
import gnlm


def nlm3d(u, h, r_p, r_s):
    return gnlm.nlm3d_autograd(u, h, r_p, r_s)
