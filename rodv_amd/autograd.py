"""torch.autograd.Function over the HIP adjoint (rodv_phmm_adjoint).

    theta = torch.tensor([45., 10.] + [1.] * 8, dtype=torch.float64, requires_grad=True)
    ll = PairHMMLogLik.apply(theta, engine, batch)     # log10 L per pair
    (-ll.sum()).backward()                              # theta.grad

theta = (gap-open Phred, gap-continuation Phred, 8 per-quality-bin scales of
the base error probability); see include/rodv_phmm.h for the model.
"""
from __future__ import annotations

import numpy as np
import torch


class PairHMMLogLik(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, engine, batch, ckpt_stride=32):
        th = theta.detach().cpu().double().numpy()
        out, _ = engine.adjoint(batch, th, weights=np.zeros(batch.num_pairs()),
                                ckpt_stride=ckpt_stride)
        ctx.engine, ctx.batch, ctx.th, ctx.ckpt = engine, batch, th, ckpt_stride
        ctx.device, ctx.dtype = theta.device, theta.dtype
        return torch.from_numpy(out).to(theta.device)

    @staticmethod
    def backward(ctx, grad_out):
        w = grad_out.detach().cpu().double().numpy()
        _, g = ctx.engine.adjoint(ctx.batch, ctx.th, weights=w, ckpt_stride=ctx.ckpt)
        return torch.from_numpy(g).to(device=ctx.device, dtype=ctx.dtype), None, None, None
