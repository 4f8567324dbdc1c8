"""torch.autograd.Function over the loss adjoint kernel (ctc_grad).

loss(F_0..F_{N-1}) = 0.5 * sum_{train} (x - <F_0(i0,:), ..., F_{N-1}(i_{N-1},:)>)^2
                   + 0.5 * lambda * sum_m sum_i c_i ||F_m(i,:)||^2
with c_i = 1 (per-row L2) or c_i = nonzeros of row i (per-occurrence, the
summed step of the reference's SGD kernel, SG/sgd.cu:92-99).

One adjoint pass per mode (a residual-weighted MTTKRP over that mode's CISS)
produces the loss and every factor gradient at once, so forward() runs the
kernel and backward() only scales the saved gradients.  The factors are the
caller's torch tensors (float32, contiguous, on the model's GPU); kernels run
on torch's current stream.
"""
from __future__ import annotations

import torch

from . import Model


class CompletionLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, model: Model, lam: float, per_occurrence: bool, *factors):
        fs = [f.contiguous() for f in factors]
        for m, f in enumerate(fs):
            if f.dtype != torch.float32 or not f.is_cuda or tuple(f.shape) != (model.dims[m], model.rank):
                raise ValueError(f"factor {m}: expected float32 cuda tensor of shape "
                                 f"({model.dims[m]}, {model.rank}), got {f.dtype} {tuple(f.shape)}")
        grads = [torch.empty_like(f) for f in fs]
        stream = torch.cuda.current_stream(fs[0].device).cuda_stream
        loss = model.grad_device([f.data_ptr() for f in fs], [g.data_ptr() for g in grads], float(lam),
                                 bool(per_occurrence), stream=stream)
        ctx.save_for_backward(*grads)
        return torch.tensor(loss, dtype=torch.float64, device=fs[0].device)

    @staticmethod
    def backward(ctx, gout):
        grads = ctx.saved_tensors
        scale = gout.to(torch.float32)
        return (None, None, None, *[g * scale for g in grads])


def completion_loss(model: Model, factors, lam: float, per_occurrence: bool = False):
    """Differentiable completion loss of `factors` (list of torch tensors)."""
    return CompletionLoss.apply(model, lam, per_occurrence, *factors)
