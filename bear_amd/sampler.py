"""Batch-capable sampler interface: live points of a nested sampler's unit
hypercube to physical parameters to log-likelihoods, many points per call.

The reference's sampler callback evaluates one point per call: MultiNest
calls Retrieval::multinestLogLikeGPU (multinest_loglike.cpp:131-186), which
maps the cube through the priors (Retrieval::convertHypercubeParameters,
multinest_loglike.cpp:59-97) and calls logLikelihoodGPU; pyBeAR's pyMultiNest
loop does the same from Python (python/emission/pyBear_retrieval_3.py:
184-202).  Here the same mapping is vectorised over a batch of cubes and the
batch goes to Retrieval.computeLikelihoodBatch (one launch sequence), so a
sampler that proposes several points at once (or several samplers, or
importance re-weighting) is fed at the batched engine's rate.

The prior transforms restate prior_types.h:39-250: uniform, log-uniform,
Gaussian (through the inverse error function), delta and linked priors, with
the optional units of BasicPrior::setUnit.  As in the reference, a linked
prior returns the value of the prior it is linked to without that prior's
unit (LinkedPrior keeps BasicPrior's unit of 1).
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

# BasicPrior::setUnit (prior_types.h:57-82), cgs (physical_const.h:40-48)
UNITS = {
    "": 1.0,
    "Me": 5.97219e27, "Mearth": 5.97219e27,
    "Mj": 1.89813e30, "Mjupiter": 1.89813e30,
    "Ms": 1.9884e33, "Msun": 1.9884e33,
    "Re": 6.37123e8, "Rearth": 6.37123e8,
    "Rj": 7.1492e9, "Rjupiter": 7.1492e9,
    "Rs": 6.96342e10, "Rsun": 6.96342e10,
    "ly": 9.4607304725808e17,
    "pc": 3.08567758135e18,
}


@dataclass
class Prior:
    """One parameter's prior: type "uniform" (bounds a, b), "log_uniform"
    (bounds a, b), "gaussian" (mean a, standard deviation b), "delta"
    (value a) or "linked" (the prior of parameter index a); unit as in the
    reference's prior files."""
    type: str
    a: float = 0.0
    b: float = 0.0
    unit: str = ""

    def __post_init__(self):
        if self.type not in ("uniform", "log_uniform", "gaussian", "delta", "linked"):
            raise ValueError(f"unknown prior type {self.type}")
        if self.unit not in UNITS:
            raise ValueError(f"Unit {self.unit} unknown")
        if self.type == "linked" and self.unit:
            raise ValueError("a linked prior takes the unit of the prior it is linked to")


def _values(prior, u):
    """parameterValue of the prior (prior_types.h:110-245) on hypercube values u."""
    if prior.type == "uniform":
        lo, hi = min(prior.a, prior.b), max(prior.a, prior.b)
        return lo + u * (hi - lo)
    if prior.type == "log_uniform":
        lo, hi = math.log10(prior.a), math.log10(prior.b)
        lo, hi = min(lo, hi), max(lo, hi)
        return 10.0 ** (lo + u * (hi - lo))
    if prior.type == "gaussian":
        from scipy.special import erfinv
        return prior.a + prior.b * math.sqrt(2.0) * erfinv(2.0 * u - 1.0)
    return np.full_like(u, prior.a, dtype=np.float64)


class Priors:
    """The priors of a retrieval's parameters, in parameter order
    (Priors::init, priors.cpp:98-131; links set up by setupLinkedPriors,
    priors.cpp:264-296)."""

    def __init__(self, priors):
        self.priors = list(priors)
        for i, p in enumerate(self.priors):
            if p.type == "linked":
                k = int(p.a)
                if not (0 <= k < len(self.priors)) or k == i:
                    raise ValueError(f"Linked prior {i}: wrong index for link")
                if self.priors[k].type == "linked":
                    raise ValueError(f"Linked prior {i}: can not link prior to another linked prior")

    def __len__(self):
        return len(self.priors)

    def transform(self, cubes):
        """Retrieval::convertHypercubeParameters for a batch: cubes [B, n] in
        the unit hypercube -> (parameters [B, n] in the prior's own units,
        physical parameters [B, n] in cgs)."""
        u = np.atleast_2d(np.asarray(cubes, dtype=np.float64))
        if u.shape[1] < len(self.priors):
            raise ValueError(f"expected {len(self.priors)} hypercube values per point, got {u.shape[1]}")
        par = np.empty((u.shape[0], len(self.priors)))
        phys = np.empty_like(par)
        for i, p in enumerate(self.priors):
            if p.type == "linked":
                k = int(p.a)
                par[:, i] = _values(self.priors[k], u[:, k])
                phys[:, i] = par[:, i]                       # the linked prior's own unit: 1
            else:
                par[:, i] = _values(p, u[:, i])
                phys[:, i] = par[:, i] * UNITS[p.unit]
        return par, phys

    def jacobian(self, cubes):
        """d(physical parameter i) / d(hypercube value source[i]) for cubes
        [B, n]: (source [n], derivatives [B, n]); source[i] is i, or the
        parameter a linked prior follows."""
        u = np.atleast_2d(np.asarray(cubes, dtype=np.float64))
        par, _ = self.transform(u)
        source = np.arange(len(self.priors))
        d = np.zeros_like(par)
        for i, p in enumerate(self.priors):
            q, scale = p, UNITS[p.unit]
            if p.type == "linked":
                source[i] = int(p.a)
                q, scale = self.priors[int(p.a)], 1.0
            if q.type == "uniform":
                d[:, i] = abs(q.b - q.a)
            elif q.type == "log_uniform":
                d[:, i] = par[:, i] * math.log(10.0) * abs(math.log10(q.b) - math.log10(q.a))
            elif q.type == "gaussian":
                d[:, i] = q.b * math.sqrt(2.0 * math.pi) * np.exp(((par[:, i] - q.a) / q.b) ** 2 / 2.0)
            d[:, i] *= scale
        return source, d


class SamplerLikelihood:
    """Log-likelihood callbacks of a nested sampler over a Retrieval.

    multinest(cube, ndim, nparams): pyMultiNest's one-point callback
    (pyBear_retrieval_3.py:184-202 with its prior function folded in): the
    cube is overwritten with the parameters, as MultiNest expects
    (multinest_loglike.cpp:84-85), and one evaluation is made.

    batch(cubes): the same for cubes [B, n] in one launch sequence; returns
    the log-likelihoods [B] and the parameters [B, n] (the values a
    one-point callback would write back into each cube).

    gradient(cubes): log-likelihoods and d logL / d(hypercube) [B, n] through
    the hand-written adjoint and the prior transforms' derivatives (for
    gradient-based samplers).
    """

    def __init__(self, retrieval, priors: Priors):
        if len(priors) != retrieval.nb_parameters:
            raise ValueError(f"{len(priors)} priors for {retrieval.nb_parameters} parameters")
        self.retrieval = retrieval
        self.priors = priors
        self.evaluations = 0

    def multinest(self, cube, ndim, nparams):
        par, phys = self.priors.transform(np.array([cube[i] for i in range(nparams)])[None, :])
        for i in range(nparams):
            cube[i] = par[0, i]
        self.evaluations += 1
        return self.retrieval.computeLikelihood(phys[0])

    def batch(self, cubes):
        par, phys = self.priors.transform(cubes)
        self.evaluations += phys.shape[0]
        return self.retrieval.computeLikelihoodBatch(phys), par

    def gradient(self, cubes):
        u = np.atleast_2d(np.asarray(cubes, dtype=np.float64))
        _, phys = self.priors.transform(u)
        ll, g = self.retrieval.gradientBatch(phys)
        source, d = self.priors.jacobian(u)
        du = np.zeros_like(phys)
        for i in range(phys.shape[1]):
            du[:, source[i]] += g[:, i] * d[:, i]
        self.evaluations += phys.shape[0]
        return ll, du
