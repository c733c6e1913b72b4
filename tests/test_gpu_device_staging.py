"""Device staging of the batched calls: each live point's temperature
profile, chemistry, altitudes and table stencils formed on the device from
its parameters (the host's Atmosphere::compute and stageLevels, restated in
stage.hip), against the same batch staged on the host.

Bars: log-likelihoods within 1e-9 relative and spectra within 1e-11
relative of the host-staged batch (the two differ only in the device's and
the host's log2 and pow, a few ulps of the densities and the gravity), the
same neglect flags; and the oracle's bars (spectra 1e-5 relative, log L
1e-6 absolute) for the device-staged batch itself."""

import numpy as np
import pytest

from bear_amd import synth
from oracle import api as oracle_api

pytestmark = pytest.mark.gpu

CASES = ["tiny", "eclipse_small", "transmission_plain", "transmission_variable_gravity", "transmission_fit_mmw",
         "transmission_fit_scale_height"]


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("name", CASES)
def test_device_staging_matches_host_staging(case_factory, name, precision):
    folder, params = case_factory(name)
    print(f"[gpu] device staging {name} {precision}", flush=True)
    r = synth.open_retrieval(folder)
    r.set_table_precision(precision)
    pts = np.array(params, dtype=np.float64)
    r.set_device_staging(False)
    ll_host = r.computeLikelihoodBatch(pts)
    spec_host, _, neg_host = r.computeModelBatch(pts)
    r.set_device_staging(True)
    ll_dev = r.computeLikelihoodBatch(pts)
    spec_dev, _, neg_dev = r.computeModelBatch(pts)
    np.testing.assert_array_equal(np.asarray(neg_dev), np.asarray(neg_host))
    live = ll_host > -1e29
    np.testing.assert_array_equal(ll_dev[~live], ll_host[~live])
    np.testing.assert_allclose(ll_dev[live], ll_host[live], rtol=1e-9, atol=0)
    rel = np.max(np.abs(spec_dev[live] - spec_host[live]) / np.abs(spec_host[live]))
    print(f"[gpu] {name} {precision}: spectra {rel:.2e}, |dlogL| {np.max(np.abs(ll_dev[live] - ll_host[live])):.2e}",
          flush=True)
    assert rel <= 1e-11
    if precision == "fp64":
        o = oracle_api.EmissionOracle(folder)
        for k, p in enumerate(pts):
            ref = o.evaluate(p)
            if ref["neglect"]:
                assert ll_dev[k] == -1e30
                continue
            assert abs(ll_dev[k] - ref["loglike"]) <= 1e-6, (k, ll_dev[k], ref["loglike"])
    r.close()


def test_device_staging_flags_neglected_points(case_factory):
    """The tiny case's cold profile (below 50 K everywhere) is neglected on
    both paths; its neighbours in the batch are not."""
    folder, params = case_factory("tiny")
    r = synth.open_retrieval(folder)
    pts = np.array(params, dtype=np.float64)
    r.set_device_staging(True)
    ll = r.computeLikelihoodBatch(pts)
    r.set_device_staging(False)
    ll_host = r.computeLikelihoodBatch(pts)
    assert (ll == -1e30).any() and (ll > -1e29).any()
    np.testing.assert_array_equal(ll == -1e30, ll_host == -1e30)
    r.close()


@pytest.mark.parametrize("name", ["tiny", "transmission_variable_gravity"])
def test_device_staged_gradient_matches_host_staging(case_factory, name):
    """The gradient calls' records staged on the device give the gradient of
    the host-staged records (the chain rule computes the atmospheres on the
    host either way)."""
    folder, params = case_factory(name)
    r = synth.open_retrieval(folder)
    pts = np.array(params, dtype=np.float64)
    r.set_device_staging(False)
    ll_h, g_h = r.gradientBatch(pts)
    r.set_device_staging(True)
    ll_d, g_d = r.gradientBatch(pts)
    live = ll_h > -1e29
    np.testing.assert_array_equal(ll_d > -1e29, live)
    np.testing.assert_allclose(ll_d[live], ll_h[live], rtol=1e-9, atol=0)
    scale = np.max(np.abs(g_h[live]), axis=0) + 1e-300
    assert np.max(np.abs(g_d[live] - g_h[live]) / scale) <= 1e-8
    r.close()
