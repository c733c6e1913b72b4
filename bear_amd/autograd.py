"""The log-likelihood as a differentiable torch function.

``log_likelihood(retrieval, parameters)`` takes a [batch, nb_parameters]
tensor of live points and returns the [batch] log-likelihoods; when the
parameters require a gradient, the forward call also runs the engine's
hand-written adjoint (bear_retrieval_gradient_batch) and backward returns
grad_output[:, None] * d logL / d parameters.  Without a gradient it is the
plain batched likelihood.
"""

from __future__ import annotations

import numpy as np
import torch


class LogLikelihood(torch.autograd.Function):
    @staticmethod
    def forward(ctx, parameters, retrieval):
        p = parameters.detach().to("cpu", torch.float64).numpy()
        if parameters.requires_grad:
            ll, grad = retrieval.gradientBatch(p)
            ctx.save_for_backward(torch.from_numpy(grad).to(parameters.device, parameters.dtype))
        else:
            ll = retrieval.computeLikelihoodBatch(p)
        return torch.from_numpy(np.ascontiguousarray(ll)).to(parameters.device, parameters.dtype)

    @staticmethod
    def backward(ctx, grad_output):
        (grad,) = ctx.saved_tensors
        return grad_output[:, None] * grad, None


def log_likelihood(retrieval, parameters: torch.Tensor) -> torch.Tensor:
    """[batch] log-likelihoods of the [batch, nb_parameters] live points."""
    if parameters.dim() == 1:
        return LogLikelihood.apply(parameters[None, :], retrieval)[0]
    return LogLikelihood.apply(parameters, retrieval)
