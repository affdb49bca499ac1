"""torch.autograd wrapper of the device forward / adjoint kernels.

`CPModel` evaluates m_k = sum_r prod_n A_n[i_n(k), r] at the stored
coordinates of a `GpuContext`'s tensor (dyn_gpu_forward_dev); its backward is
the adjoint spMTTKRP of the incoming gradient vector for every mode
(dyn_gpu_adjoint_dev).  Both kernels run on torch's current stream.

`fit_completion` is the sparse-completion driver: generalised CP with the
Bernoulli-logit (or least-squares) loss, optimised with Adam on gradients from
the fused loss+adjoint kernel (dyn_gpu_completion_grad_dev).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .engine import GpuContext, padded_rank


def _torch():
    import torch
    return torch


def _ptrs(tensors):
    return (L._vp * len(tensors))(*[t.data_ptr() for t in tensors])


def _pad(A, RP):
    torch = _torch()
    return [torch.nn.functional.pad(a, (0, RP - a.shape[1])).contiguous() for a in A]


class _CPForward:
    @staticmethod
    def build():
        torch = _torch()

        class Fn(torch.autograd.Function):
            @staticmethod
            def forward(ctx, engine: GpuContext, RP: int, *A):
                m = torch.empty(engine.nnz, device=A[0].device, dtype=torch.float32)
                with engine.on_stream(torch.cuda.current_stream(A[0].device).cuda_stream):
                    engine._check(engine.lib.dyn_gpu_forward_dev(engine.h, _ptrs(A), RP,
                                                                 L._vp(m.data_ptr())),
                                  "forward_dev")
                ctx.engine = engine
                ctx.RP = RP
                ctx.save_for_backward(*A)
                return m

            @staticmethod
            def backward(ctx, gm):
                A = ctx.saved_tensors
                engine = ctx.engine
                gm = gm.contiguous().to(torch.float32)
                grads = []
                with engine.on_stream(torch.cuda.current_stream(gm.device).cuda_stream):
                    for n, a in enumerate(A):
                        if not ctx.needs_input_grad[2 + n]:
                            grads.append(None)
                            continue
                        g = torch.empty_like(a)
                        engine._check(engine.lib.dyn_gpu_adjoint_dev(
                            engine.h, L._vp(gm.data_ptr()), n, _ptrs(A), ctx.RP,
                            L._vp(g.data_ptr())), f"adjoint_dev({n})")
                        grads.append(g)
                return (None, None, *grads)

        return Fn


_FN = None


def cp_forward(engine: GpuContext, factors):
    """m = CP model at the stored coordinates (differentiable w.r.t. factors).
    `factors`: list of (I_n, R) float32 CUDA tensors."""
    global _FN
    if _FN is None:
        _FN = _CPForward.build()
    R = factors[0].shape[1]
    RP = padded_rank(R)
    return _FN.apply(engine, RP, *_pad(factors, RP))


def bernoulli_logit_loss(m, y):
    torch = _torch()
    return (torch.nn.functional.softplus(m) - y * m).sum()


def completion_loss_and_grad(engine: GpuContext, factors, loss: str = "bernoulli"):
    """Fused device loss + gradients (dyn_gpu_completion_grad_dev) with the
    stored tensor values as labels.  Returns (loss, [grad_n])."""
    torch = _torch()
    R = factors[0].shape[1]
    RP = padded_rank(R)
    A = _pad([f.detach() for f in factors], RP)
    grads = [torch.empty_like(a) for a in A]
    out = C.c_double(0.0)
    kind = L.DYN_LOSS_BERNOULLI if loss == "bernoulli" else L.DYN_LOSS_LS
    with engine.on_stream(torch.cuda.current_stream(A[0].device).cuda_stream):
        engine._check(engine.lib.dyn_gpu_completion_grad_dev(engine.h, kind, _ptrs(A), RP,
                                                             _ptrs(grads), C.byref(out)),
                      "completion_grad_dev")
    return float(out.value), [g[:, :R] for g in grads]


def fit_completion(engine: GpuContext, dims, rank: int, steps: int = 100, lr: float = 1e-2,
                   betas=(0.9, 0.999), eps: float = 1e-8, init_std: float = 0.1, seed: int = 0,
                   loss: str = "bernoulli", device: str = "cuda"):
    """Adam on the sparse-completion loss; returns (factors, per-step losses)."""
    torch = _torch()
    g = torch.Generator(device="cpu").manual_seed(seed)
    A = [(init_std * torch.randn(int(d), rank, generator=g)).to(device).requires_grad_(True)
         for d in dims]
    opt = torch.optim.Adam(A, lr=lr, betas=betas, eps=eps)
    losses = []
    for _ in range(steps):
        val, grads = completion_loss_and_grad(engine, A, loss)
        for a, gr in zip(A, grads):
            a.grad = gr.contiguous()
        opt.step()
        losses.append(val)
    return [a.detach() for a in A], np.array(losses)
