"""The batch-capable sampler interface (bear_amd/sampler.py): the prior
transforms of the reference's priors (prior_types.h:39-250) on the CPU, and
on the GPU the batched callback against one-point callbacks."""

import math

import numpy as np
import pytest

from bear_amd import sampler


def test_prior_transforms_closed_forms():
    from scipy.stats import norm
    pr = sampler.Priors([
        sampler.Prior("uniform", 5.0, 2.0),                 # bounds in either order (prior_types.h:101-102)
        sampler.Prior("log_uniform", 1e-8, 1e-2),
        sampler.Prior("gaussian", 3.5, 0.2),
        sampler.Prior("delta", 1.7, unit="Rj"),
        sampler.Prior("uniform", 0.8, 1.2, unit="Rj"),
        sampler.Prior("linked", 4),
    ])
    u = np.random.default_rng(0).uniform(0.01, 0.99, (50, 6))
    par, phys = pr.transform(u)
    np.testing.assert_allclose(par[:, 0], 2.0 + 3.0 * u[:, 0], rtol=1e-15)
    np.testing.assert_allclose(np.log10(par[:, 1]), -8.0 + 6.0 * u[:, 1], rtol=1e-13)
    np.testing.assert_allclose(par[:, 2], norm.ppf(u[:, 2], 3.5, 0.2), rtol=1e-12)
    assert np.all(par[:, 3] == 1.7) and np.all(phys[:, 3] == 1.7 * 7.1492e9)
    np.testing.assert_allclose(phys[:, 4], (0.8 + 0.4 * u[:, 4]) * 7.1492e9, rtol=1e-15)
    # a linked prior follows its source's hypercube value, without the unit
    np.testing.assert_array_equal(par[:, 5], par[:, 4])
    np.testing.assert_array_equal(phys[:, 5], par[:, 4])
    assert pr.transform(u[:1])[0].shape == (1, 6)


def test_prior_jacobian_central_differences():
    pr = sampler.Priors([sampler.Prior("uniform", 2.0, 5.0), sampler.Prior("log_uniform", 1e-8, 1e-2),
                         sampler.Prior("gaussian", 3.5, 0.2, unit="pc"), sampler.Prior("delta", 1.0),
                         sampler.Prior("linked", 1)])
    u = np.random.default_rng(1).uniform(0.05, 0.95, (20, 5))
    source, d = pr.jacobian(u)
    np.testing.assert_array_equal(source, [0, 1, 2, 3, 1])
    h = 1e-6
    for i in range(5):
        e = np.zeros(5)
        e[source[i]] = h
        fd = (pr.transform(u + e)[1][:, i] - pr.transform(u - e)[1][:, i]) / (2 * h)
        np.testing.assert_allclose(d[:, i], fd, rtol=1e-6, atol=1e-12 * np.abs(d[:, i]).max() + 1e-300)


def test_prior_errors():
    with pytest.raises(ValueError):
        sampler.Prior("beta", 1.0, 2.0)
    with pytest.raises(ValueError):
        sampler.Prior("uniform", 1.0, 2.0, unit="furlong")
    with pytest.raises(ValueError):
        sampler.Priors([sampler.Prior("uniform", 0, 1), sampler.Prior("linked", 1)])
    with pytest.raises(ValueError):
        sampler.Priors([sampler.Prior("uniform", 0, 1), sampler.Prior("linked", 0), sampler.Prior("linked", 1)])


def _emission_priors(case):
    """Priors around a case's parameters: log g uniform, f and d delta, the
    mixing ratios log-uniform, the temperature DOF uniform, shifts Gaussian."""
    p = case.parameters
    priors = [sampler.Prior("uniform", p[0] - 0.2, p[0] + 0.2), sampler.Prior("delta", p[1]),
              sampler.Prior("delta", p[2] / 3.08567758135e18, unit="pc")]
    nb_mix = sum(len(c[1]) for c in case.chemistry if c[0] == "iso")
    priors += [sampler.Prior("log_uniform", p[3 + k] / 3, p[3 + k] * 3) for k in range(nb_mix)]
    priors += [sampler.Prior("uniform", p[3 + nb_mix] * 0.95, p[3 + nb_mix] * 1.05)]
    priors += [sampler.Prior("uniform", 0.97 * v, 1.03 * v) for v in p[4 + nb_mix:10 + nb_mix]]
    priors += [sampler.Prior("gaussian", 0.0, 1e-17) for _ in p[10 + nb_mix:]]
    return sampler.Priors(priors)


@pytest.mark.gpu
def test_batched_callback_equals_one_point_callbacks(case_factory):
    """B hypercube points through the batched callback give what B calls of
    the one-point (pyMultiNest) callback give, and write back the same
    parameters."""
    from bear_amd import synth
    folder, _ = case_factory("tiny")
    r = synth.open_retrieval(folder)
    case = synth.tiny_case()
    like = sampler.SamplerLikelihood(r, _emission_priors(case))
    cubes = np.random.default_rng(3).uniform(0.02, 0.98, (24, r.nb_parameters))
    ll, par = like.batch(cubes)
    for k in range(cubes.shape[0]):
        cube = list(cubes[k])
        one = like.multinest(cube, r.nb_parameters, r.nb_parameters)
        np.testing.assert_array_equal(np.array(cube), par[k])
        assert abs(one - ll[k]) <= 1e-9 * max(1.0, abs(one)), (k, one, ll[k])
    assert like.evaluations == 2 * cubes.shape[0]
    print(f"[gpu] sampler: {cubes.shape[0]} points, log L in [{ll.min():.3f}, {ll.max():.3f}]", flush=True)
    r.close()


@pytest.mark.gpu
def test_hypercube_gradient(case_factory):
    """d logL / d(hypercube) from the adjoint and the priors' derivatives
    against central differences of the batched callback at steps h and h/2
    (Richardson-combined), within 2e-4 relative plus ten times the
    differences' own error estimate |D(h) - D(h/2)| (the bilinear table
    interpolation has kinks at the table nodes)."""
    from bear_amd import synth
    folder, _ = case_factory("tiny")
    r = synth.open_retrieval(folder)
    like = sampler.SamplerLikelihood(r, _emission_priors(synth.tiny_case()))
    u = np.random.default_rng(4).uniform(0.2, 0.8, (2, r.nb_parameters))
    ll, du = like.gradient(u)
    np.testing.assert_allclose(ll, like.batch(u)[0], rtol=1e-12)
    scale = np.abs(du).max()
    for i in range(u.shape[1]):
        d = []
        for h in (1e-5, 5e-6):
            e = np.zeros_like(u)
            e[:, i] = h
            d.append((like.batch(u + e)[0] - like.batch(u - e)[0]) / (2 * h))
        fd = (4 * d[1] - d[0]) / 3
        tol = 2e-4 * np.abs(fd) + 10 * np.abs(d[1] - d[0]) + 1e-6 * scale
        print(f"[gpu] hypercube gradient {i}: adjoint {du[:, i]}, differences {fd}", flush=True)
        assert np.all(np.abs(du[:, i] - fd) <= tol), i
    r.close()
