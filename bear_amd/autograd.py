"""The log-likelihood as a differentiable torch function.

``log_likelihood(retrieval, parameters)`` takes a [batch, nb_parameters]
tensor of live points and returns the [batch] log-likelihoods; backward
returns grad_output[:, None] * d logL / d parameters from the engine's
hand-written adjoint (bear_retrieval_gradient_batch).  Without a gradient it
is the plain batched likelihood.

gradient="eager" (the default) runs value and gradient together in the
forward call when the parameters require a gradient: one launch sequence,
the cheapest choice when backward follows (gradient-based samplers,
optimisers).  gradient="lazy" runs only the batched likelihood forward and
the value + gradient sequence in backward, for graphs where backward often
does not run; a backward then costs one extra forward pass.

The [batch, nb_parameters] values cross to the host once per call (a few
KB) and go back to the device with the batch's other inputs: the batched
calls form each point's atmosphere and table stencils on the device from
them (device staging, the default), and the chain rule from the level
adjoints to the parameters runs on the host in FP64 (a few microseconds per
point).  The returned log-likelihoods and gradients are copied to the
parameters' device.
"""

from __future__ import annotations

import numpy as np
import torch


def _host(parameters):
    return parameters.detach().to("cpu", torch.float64).numpy()


class LogLikelihood(torch.autograd.Function):
    @staticmethod
    def forward(ctx, parameters, retrieval, lazy):
        p = _host(parameters)
        ctx.lazy = bool(lazy)
        ctx.retrieval = retrieval
        if parameters.requires_grad and not ctx.lazy:
            ll, grad = retrieval.gradientBatch(p)
            ctx.save_for_backward(torch.from_numpy(grad).to(parameters.device, parameters.dtype))
        else:
            ll = retrieval.computeLikelihoodBatch(p)
            ctx.save_for_backward(parameters.detach())
        return torch.from_numpy(np.ascontiguousarray(ll)).to(parameters.device, parameters.dtype)

    @staticmethod
    def backward(ctx, grad_output):
        (saved,) = ctx.saved_tensors
        if ctx.lazy:
            _, grad = ctx.retrieval.gradientBatch(_host(saved))
            saved = torch.from_numpy(grad).to(grad_output.device, grad_output.dtype)
        return grad_output[:, None] * saved, None, None


def log_likelihood(retrieval, parameters: torch.Tensor, gradient: str = "eager") -> torch.Tensor:
    """[batch] log-likelihoods of the [batch, nb_parameters] live points
    (a [nb_parameters] vector gives a scalar)."""
    if gradient not in ("eager", "lazy"):
        raise ValueError(f"gradient must be 'eager' or 'lazy', not {gradient!r}")
    lazy = gradient == "lazy"
    if parameters.dim() == 1:
        return LogLikelihood.apply(parameters[None, :], retrieval, lazy)[0]
    return LogLikelihood.apply(parameters, retrieval, lazy)
