"""L2 misfit of canonical records, computed by one HIP pass.

``l2_misfit(pred, obs, mask)`` = 0.5 * sum(((pred - obs) * mask)^2), the loss
of the reference's tests and gradient checks, with padded receivers masked
out (a receiver of mask 0 contributes nothing, its adjoint source is 0).
The sum is accumulated in float64 in a fixed order, so the same inputs give
the same loss on every run (what a finite-difference check needs).  The
backward returns the stored residual (times the incoming scalar) as the
gradient of ``pred``: the adjoint source, no second pass over the records.
"""
from __future__ import annotations

import torch

from . import _core


class _L2Misfit(torch.autograd.Function):

    @staticmethod
    def forward(ctx, pred, obs, mask):
        if pred.shape != obs.shape or pred.ndim != 4:
            raise ValueError(f"pred {tuple(pred.shape)} and obs {tuple(obs.shape)} must both be canonical "
                             "(B, nt, nrec, nfield) records")
        for t, what in ((pred, "pred"), (obs, "obs")):
            if t.dtype != torch.float32 or not t.is_cuda:
                raise TypeError(f"{what} must be a float32 tensor on the GPU")
        B, nt, nrec, nf = pred.shape
        if mask is not None:
            mask = mask.to(device=pred.device, dtype=torch.float32)
            if mask.numel() != B * nrec:
                raise ValueError(f"mask must hold one weight per (shot, receiver), (B, nrec) = ({B}, {nrec}); "
                                 f"got shape {tuple(mask.shape)}")
            mask = mask.reshape(B, nrec).contiguous()
        pred = pred.contiguous()
        obs = obs.contiguous()
        need_res = ctx.needs_input_grad[0]
        res = torch.empty_like(pred) if need_res else None
        loss = torch.empty((), dtype=torch.float64, device=pred.device)
        partials = torch.empty(_core.MISFIT_PARTIALS, dtype=torch.float64, device=pred.device)
        L = _core.load()
        _core.on_stream(pred.device, L.swp_l2_misfit, pred.data_ptr(), obs.data_ptr(),
                        mask.data_ptr() if mask is not None else None, nf, B, nrec, nt,
                        res.data_ptr() if res is not None else None, loss.data_ptr(), partials.data_ptr())
        ctx.save_for_backward(res if res is not None else pred.new_empty(0))
        return loss

    @staticmethod
    def backward(ctx, g):
        (res,) = ctx.saved_tensors
        if not ctx.needs_input_grad[0]:
            return None, None, None
        return res * g.to(torch.float32), None, None


def l2_misfit(pred: torch.Tensor, obs: torch.Tensor, mask: torch.Tensor | None = None) -> torch.Tensor:
    """0.5 * ||(pred - obs) * mask||^2 as a float64 scalar tensor.

    pred, obs: canonical ``(B, nt, nrec, nfield)`` float32 records on the GPU;
    mask: ``(B, nrec)`` weights (any shape with B * nrec elements) or None."""
    return _L2Misfit.apply(pred, obs, mask)
