"""The ring-polymer action as a differentiable torch function.

``ring_action(x, ring)`` returns U_M of every chain of a ``Ring`` (instantonmod.f90:17-151)
for the concatenated chain coordinates ``x`` (a float64 tensor, chains in the reference
layout (n, ndim, natom) column-major, one after another).  Its backward pass is the
device adjoint (rp_ring_um), and the backward of that backward is the device
forward-over-reverse Hessian-vector product (rp_ring_hvp), so first and second
derivatives of any scalar built from the actions come from the hand-written kernels.
Tensors may live on any device; the library takes host buffers and copies them itself.

The Ring is the function's workspace: every evaluation loads x into its device coordinates
and leaves them there (keep an optimised instanton in a Ring of its own).  The gradient made
by the forward pass is kept for the backward pass, so a first derivative costs one adjoint
evaluation in all.
"""
from __future__ import annotations

import numpy as np
import torch

from .instanton import Ring


def _host(t: torch.Tensor) -> np.ndarray:
    return np.ascontiguousarray(t.detach().to("cpu", torch.float64).numpy())


class _Action(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, ring):
        flat = _host(x)
        ring.set_x(ring.unpack(flat))
        g, e = ring.UMforceenergy()
        ctx.ring = ring
        ctx.g = g
        ctx.save_for_backward(x)
        return torch.as_tensor(e, dtype=x.dtype, device=x.device)

    @staticmethod
    def backward(ctx, gout):
        (x,) = ctx.saved_tensors
        return _ActionGrad.apply(x, gout, ctx.ring, ctx.g), None


class _ActionGrad(torch.autograd.Function):
    """(x, w) -> sum_c w_c grad U_c(x), laid out like x; g = the gradients at x when the
    caller already holds them."""

    @staticmethod
    def forward(ctx, x, w, ring, g=None):
        if g is None:
            ring.set_x(ring.unpack(_host(x)))
            g, _ = ring.UMforceenergy()
        wh = _host(w)
        out = np.concatenate([wh[c] * g[c].reshape(-1) for c in range(ring.nchain)])
        ctx.ring = ring
        ctx.save_for_backward(x, w)
        ctx.g = g
        return torch.as_tensor(out, dtype=x.dtype, device=x.device)

    @staticmethod
    def backward(ctx, gg):
        x, w = ctx.saved_tensors
        ring = ctx.ring
        ggh = _host(gg)
        wh = _host(w)
        # d/dx: sum_c w_c H_c gg_c (the HVP runs at the x of the forward pass)
        ring.set_x(ring.unpack(_host(x)))
        hv = ring.hvp(ring.unpack(ggh))
        dx = np.concatenate([wh[c] * hv[c].reshape(-1) for c in range(ring.nchain)])
        # d/dw_c: grad U_c . gg_c
        vs = ring.unpack(ggh)
        dw = np.array([float(np.dot(ctx.g[c].reshape(-1), vs[c].reshape(-1)))
                       for c in range(ring.nchain)])
        return (torch.as_tensor(dx, dtype=x.dtype, device=x.device),
                torch.as_tensor(dw, dtype=w.dtype, device=w.device), None, None)


def ring_action(x: torch.Tensor, ring: Ring) -> torch.Tensor:
    """U_M of every chain of ``ring`` at ``x`` (shape (sum_c n_c * ndof,)), differentiable
    twice: gradients from the adjoint kernel, second derivatives from its HVP."""
    return _Action.apply(x, ring)
