"""torch.autograd binding of the accurate adjoint NFFT (thin test binding).

    F = adjoint_nfft(x, y, M)         x, y: (n,) or (B, n) float tensors on the GPU
    F[..., k + M/2] = sum_j y_j exp(2 pi i k x_j),  k = -M/2 .. M/2-1

Forward and backward run the library's HIP kernels: the backward is the
forward (type-2) NFFT of the output gradient gathered with the same Gaussian
weights (d/dy) and with their derivative (d/dx); the weights are recomputed
from the plan, not stored.
"""
from __future__ import annotations

import numpy as np
import torch

from .api import NFFTPlan


class AdjointNFFTFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, y, M: int, sigma: int, m, tile_log2: int):
        batched = x.dim() == 2
        xs = x.detach().contiguous()
        ys = y.detach().contiguous()
        if batched:
            B, n = xs.shape
            offsets = np.arange(B + 1, dtype=np.int32) * n
            plan = NFFTPlan(xs.reshape(-1), M, offsets=offsets, sigma=sigma, m=m,
                            tile_log2=tile_log2)
            out = plan.adjoint(ys.reshape(-1))
        else:
            plan = NFFTPlan(xs, M, sigma=sigma, m=m, tile_log2=tile_log2)
            out = plan.adjoint(ys)[0]
        ctx.plan = plan
        ctx.batched = batched
        ctx.shape = x.shape
        ctx.save_for_backward(ys)
        return out

    @staticmethod
    def backward(ctx, grad_out):
        (ys,) = ctx.saved_tensors
        plan = ctx.plan
        want_x, want_y = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
        g = grad_out.contiguous()
        if not ctx.batched:
            g = g.reshape(1, -1)
        dy, dx = plan.backward(ys.reshape(-1), g, want_y=want_y, want_x=want_x)
        dx = dx.reshape(ctx.shape) if dx is not None else None
        dy = dy.reshape(ctx.shape) if dy is not None else None
        return dx, dy, None, None, None, None


def adjoint_nfft(x, y, M: int, sigma: int = 2, m: int | None = None, tile_log2: int = 0):
    return AdjointNFFTFunction.apply(x, y, M, sigma, m, tile_log2)
