"""The ring-polymer action as a differentiable torch function.

``ring_action(x, ring)`` returns U_M of every chain of a ``Ring`` (instantonmod.f90:17-151)
for the concatenated chain coordinates ``x`` (a float64 tensor, chains in the reference
layout (n, ndim, natom) column-major, one after another).  Its backward pass is the
device adjoint (rp_ring_um), and the backward of that backward is the device
forward-over-reverse Hessian-vector product (rp_ring_hvp), so first and second
derivatives of any scalar built from the actions come from the hand-written kernels.
A tensor on the GPU stays there: the library's device-memory entry points
(rp_ring_set_x_device, rp_ring_um_device, rp_ring_hvp_device) copy it device to device, and
the per-chain weighting is done by torch on the same device.  CPU tensors go through the
host entry points.

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


def _device(t: torch.Tensor) -> torch.Tensor:
    """t as a contiguous float64 tensor on its GPU, every write to it complete."""
    d = t.detach().to(torch.float64).contiguous()
    torch.cuda.current_stream(d.device).synchronize()
    return d


def _flat(chains) -> torch.Tensor:
    return torch.from_numpy(np.concatenate([c.reshape(-1) for c in chains]))


def _load_x(ring: Ring, x: torch.Tensor) -> None:
    if x.is_cuda:
        ring.set_x_device(_device(x).data_ptr())
    else:
        ring.set_x(ring.unpack(_host(x)))


def _um(ring: Ring, x: torch.Tensor):
    """(U_M per chain as a host array, the flat adjoint gradient on x's device, float64)"""
    _load_x(ring, x)
    if x.is_cuda:
        g = torch.empty(x.numel(), dtype=torch.float64, device=x.device)
        return ring.um_device(g.data_ptr()), g
    g, e = ring.UMforceenergy()
    return e, _flat(g)


def _hvp(ring: Ring, x: torch.Tensor, v: torch.Tensor) -> torch.Tensor:
    """H(x) v, flat, on x's device (float64)"""
    _load_x(ring, x)
    if x.is_cuda:
        vd = _device(v)
        hv = torch.empty_like(vd)
        ring.hvp_device(vd.data_ptr(), hv.data_ptr())
        return hv
    return _flat(ring.hvp(ring.unpack(_host(v))))


def _chain_index(ring: Ring, device) -> torch.Tensor:
    """the chain of every element of a flat coordinate vector"""
    sizes = torch.as_tensor(np.diff(ring.off), dtype=torch.int64)
    return torch.repeat_interleave(torch.arange(ring.nchain), sizes).to(device)


class _Action(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, ring):
        e, g = _um(ring, x)
        ctx.ring = ring
        ctx.g = g
        ctx.save_for_backward(x)
        return torch.as_tensor(e, dtype=x.dtype, device=x.device)

    @staticmethod
    def backward(ctx, gout):
        (x,) = ctx.saved_tensors
        return _ActionGrad.apply(x, gout, ctx.ring, ctx.g), None


class _ActionGrad(torch.autograd.Function):
    """(x, w) -> sum_c w_c grad U_c(x), laid out like x; g = the flat gradient at x (float64,
    on x's device) when the caller already holds it."""

    @staticmethod
    def forward(ctx, x, w, ring, g=None):
        if g is None:
            _, g = _um(ring, x)
        cix = _chain_index(ring, x.device)
        ctx.ring = ring
        ctx.save_for_backward(x, w)
        ctx.g = g
        ctx.cix = cix
        return (w.to(torch.float64)[cix] * g).to(x.dtype)

    @staticmethod
    def backward(ctx, gg):
        x, w = ctx.saved_tensors
        ring, cix = ctx.ring, ctx.cix
        # d/dx: sum_c w_c H_c gg_c (the HVP runs at the x of the forward pass)
        hv = _hvp(ring, x, gg)
        dx = w.to(torch.float64)[cix] * hv
        # d/dw_c: grad U_c . gg_c
        dw = torch.zeros(ring.nchain, dtype=torch.float64, device=x.device)
        dw.index_add_(0, cix, ctx.g * gg.to(torch.float64))
        return dx.to(x.dtype), dw.to(w.dtype), None, None


def ring_action(x: torch.Tensor, ring: Ring) -> torch.Tensor:
    """U_M of every chain of ``ring`` at ``x`` (shape (sum_c n_c * ndof,)), differentiable
    twice: gradients from the adjoint kernel, second derivatives from its HVP."""
    return _Action.apply(x, ring)
