"""Parity of the HIP transmission and secondary-eclipse paths with the CPU
oracle, through the C ABI: spectra and bins within 1e-5 relative,
log-likelihood within 1e-6 absolute."""

import numpy as np
import pytest

from bear_amd import synth
from oracle import api as oracle_api

pytestmark = pytest.mark.gpu

SPECTRUM_RTOL = 1e-5
LOGLIKE_ATOL = 1e-6

MODELS = ["eclipse_small", "eclipse_wide", "transmission_plain", "transmission_variable_gravity",
          "transmission_fit_mmw", "transmission_fit_scale_height", "transmission_hminus",
          "cloudy_eclipse", "transmission_cloudy"]


@pytest.mark.parametrize("name", MODELS)
def test_model_parity(case_factory, name):
    folder, params = case_factory(name)
    print(f"[gpu] {name} at {folder}", flush=True)
    r = synth.open_retrieval(folder)
    o = oracle_api.EmissionOracle(folder)
    assert r.nb_parameters == o.nb_parameters
    nu, _ = r.spectral_grid()
    np.testing.assert_array_equal(nu, o.grid()[0])
    for k, p in enumerate(params):
        print(f"[gpu] {name} parameter vector {k}", flush=True)
        ref = o.evaluate(p)
        out = r.computeModel(p, True)
        ll = r.computeLikelihood(p)
        assert out.neglect_model == ref["neglect"]
        rel = np.max(np.abs(out.spectrum - ref["spectrum"]) / np.abs(ref["spectrum"]))
        assert rel <= SPECTRUM_RTOL, (k, rel)
        for i in range(o.nb_observations):
            ob = o.observation(i)
            expect = ref["bins"][i][::-1] if ob["ascending"] else ref["bins"][i]
            np.testing.assert_allclose(out.spectrum_obs[i], expect, rtol=SPECTRUM_RTOL, atol=0)
        assert abs(ll - ref["loglike"]) <= LOGLIKE_ATOL, (k, ll, ref["loglike"])
    r.close()


def test_transparent_transmission_on_gpu(case_factory):
    """Zero abundances and no scatterer: depth = (R_p / R_*)^2 1e6 at every wavenumber."""
    folder, params = case_factory("transmission_plain")
    r = synth.open_retrieval(folder)
    p = params[0].copy()
    p[3] = p[4] = 0.0
    out = r.computeModel(p, True)
    # CIA (n_H2^2) and H2 Rayleigh remain: the depth is at least the bare disc
    assert np.all(out.spectrum >= (p[1] / p[2]) ** 2 * 1e6 * (1 - 1e-12))
    r.close()
