"""Thin torch.autograd binding of the hand-written adjoints, used to check them.

``SpectralFilter.apply(u, coeffs, f)`` computes v = C2R(H(coeffs) * R2C(u)) with a
radial shell filter H on an initialised fused ``ffts`` instance ``f`` (one
``ffts.filter`` call).  Its backward
uses only the C-ABI adjoint kernels and recomputes the forward R2C instead of
storing the spectrum:

    X      = R2C(u)                       (recomputed)
    G      = C2R^T(dL/dv)                 (dL/dZ, Z = H X)
    dL/dc  = shell_gradient(G, X)
    dL/du  = R2C^T(H G)
"""

from __future__ import annotations

import torch


class SpectralFilter(torch.autograd.Function):

    @staticmethod
    def forward(ctx, u, coeffs, f):
        # one fused filter call (the spectrum stays in the work area); bitwise equal to
        # forward + radial multiply + backward
        v = f.real_array()
        f.filter(u.contiguous(), v, "shell", scale=1.0, coeffs=coeffs.contiguous())
        ctx.f = f
        ctx.save_for_backward(u, coeffs)
        return v

    @staticmethod
    def backward(ctx, grad_v):
        u, coeffs = ctx.saved_tensors
        f = ctx.f
        X = f.complex_array()
        f.forward(u.contiguous(), X)
        G = f.complex_array()
        f.c2r_adjoint(grad_v.contiguous(), G)
        grad_c = f.shell_gradient(G, X, coeffs.numel())
        f.spectral_multiply(G, G, shell_coeffs=coeffs.contiguous())
        grad_u = f.real_array()
        f.r2c_adjoint(G, grad_u)
        return grad_u, grad_c, None


def spectral_filter(u, coeffs, f):
    return SpectralFilter.apply(u, coeffs, f)
