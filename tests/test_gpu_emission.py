"""Parity of the HIP emission path with the CPU oracle, through the C ABI.

Bars (stated by the task for the likelihood path): high-resolution spectra
and bins within 1e-5 relative of the oracle, log-likelihood within 1e-6
absolute."""

import json
import math
import os

import numpy as np
import pytest

from bear_amd import api, synth
from oracle import api as oracle_api

pytestmark = pytest.mark.gpu

SPECTRUM_RTOL = 1e-5
LOGLIKE_ATOL = 1e-6

_RETRIEVALS = {}


def retrieval(case_factory, name):
    if name not in _RETRIEVALS:
        folder, params = case_factory(name)
        print(f"[gpu] opening {name} at {folder}", flush=True)
        _RETRIEVALS[name] = (synth.open_retrieval(folder), oracle_api.EmissionOracle(folder), params, folder)
    return _RETRIEVALS[name]


@pytest.mark.parametrize("name", ["tiny", "emission_c1", "tiny_const_wavelength", "tiny_const_resolution"])
def test_grid_matches_oracle(case_factory, name):
    r, o, _, _ = retrieval(case_factory, name)
    nu, lam = r.spectral_grid()
    onu, olam, _ = o.grid()
    np.testing.assert_array_equal(nu, onu)
    np.testing.assert_array_equal(lam, olam)
    assert r.nb_parameters == o.nb_parameters
    assert r.nb_observation_points == o.nb_bins


@pytest.mark.parametrize("name", ["tiny", "emission_c1", "emission_wide", "tiny_const_wavelength",
                                  "tiny_const_resolution", "many_species", "hminus", "cloudy_tiny"])
def test_emission_parity(case_factory, name):
    r, o, params, _ = retrieval(case_factory, name)
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
        if ref["neglect"]:
            assert ll == -1e30
        else:
            assert abs(ll - ref["loglike"]) <= LOGLIKE_ATOL, (k, ll, ref["loglike"])


def test_isothermal_identity_on_gpu(case_factory):
    """pi B(T) 1e-3 scale at every wavenumber, independently of the tables."""
    r, o, params, _ = retrieval(case_factory, "emission_c1")
    p = params[0].copy()
    p[3 + 2 + 1:] = 1.0
    out = r.computeModel(p, True)
    nu, lam = r.spectral_grid()
    h, c, kb = synth.PLANCK_H, synth.LIGHT_C, synth.BOLTZMANN_K
    o.evaluate(p)
    t = o.atmosphere()["temperature"][0]
    b = 2 * h * c * c * nu ** 3 / np.expm1(h * c * nu / (kb * t))
    scale = (synth.RADIUS_JUPITER / p[2]) ** 2 * p[1]
    expected = math.pi * b * 1e-3 * scale / lam ** 2 * 1e4
    np.testing.assert_allclose(out.spectrum, expected, rtol=1e-10)


def test_in_memory_inputs_match_files(case_factory, tmp_path):
    """Tables registered in memory + ObservationInput == the file-based set-up."""
    name = "tiny"
    r, _, params, folder = retrieval(case_factory, name)
    case = synth.tiny_case()
    meta = json.load(open(os.path.join(folder, "case.json")))
    # the injected data come from the written files: read them back through the oracle
    o = oracle_api.EmissionOracle(folder)
    for i, spec in enumerate(case.observations):
        ob = o.observation(i)
        spec.data = ob["data"][::-1].copy() if ob["ascending"] else ob["data"]
        spec.error = ob["error"][::-1].copy() if ob["ascending"] else ob["error"]
    mem = synth.open_retrieval_in_memory(case, str(tmp_path / "scratch"))
    for p in params[:3]:
        a, b = r.computeModel(p, True), mem.computeModel(p, True)
        np.testing.assert_array_equal(a.spectrum, b.spectrum)
        for x, y in zip(a.spectrum_obs, b.spectrum_obs):
            np.testing.assert_array_equal(x, y)
        assert r.computeLikelihood(p) == mem.computeLikelihood(p)
    mem.close()
    assert meta["parameters"]


def test_calc_model_gpu_into_torch_buffers(case_factory):
    import torch
    r, o, params, _ = retrieval(case_factory, "tiny")
    p = params[1]
    spec = torch.zeros(r.nb_spectral_points, dtype=torch.float64, device="cuda")
    obs = [torch.zeros(n, dtype=torch.float64, device="cuda") for n in r.nb_observation_points]
    torch.cuda.synchronize()
    neglect = r.calcModelGPU(p, spec.data_ptr(), [t.data_ptr() for t in obs])
    ref = o.evaluate(p)
    assert not neglect
    np.testing.assert_allclose(spec.cpu().numpy(), ref["spectrum"], rtol=SPECTRUM_RTOL)
    for i, t in enumerate(obs):           # device buffers hold the internal (descending) order
        np.testing.assert_allclose(t.cpu().numpy(), ref["bins"][i], rtol=SPECTRUM_RTOL)


def test_repeated_evaluations_are_deterministic(case_factory):
    r, _, params, _ = retrieval(case_factory, "emission_c1")
    lls = [r.computeLikelihood(params[1]) for _ in range(5)]
    assert len(set(lls)) == 1


def test_first_evaluation_after_setup(case_factory):
    """The first evaluation of a fresh retrieval sees the uploaded tables
    (set-up copies and kernels share one stream, and set-up waits for them)."""
    folder, params = case_factory("emission_c1")
    o = oracle_api.EmissionOracle(folder)
    for k in range(2):
        r = synth.open_retrieval(folder)
        out = r.computeModel(params[1], True)          # first call on this retrieval
        ref = o.evaluate(params[1])
        rel = np.max(np.abs(out.spectrum - ref["spectrum"]) / np.abs(ref["spectrum"]))
        assert rel <= SPECTRUM_RTOL, (k, rel)
        r.close()


def test_many_species_limits(case_factory):
    """22 tabulated species run the generic single-point kernel (parity in
    test_emission_parity); the register-cached batched and adjoint kernels
    hold at most 16 stencil slots and refuse the case with InvalidInput."""
    r, o, params, _ = retrieval(case_factory, "many_species")
    assert r.stats()["nb_species"] == 22
    pts = np.array(params[:2])
    with pytest.raises(api.BearError) as e:
        r.computeLikelihoodBatch(pts)
    assert e.value.code == -1 and "16" in str(e.value)
    with pytest.raises(api.BearError) as e:
        r.gradientBatch(pts)
    assert e.value.code == -1
    # the single-point path still works after the refusals
    assert abs(r.computeLikelihood(params[0]) - o.evaluate(params[0])["loglike"]) <= LOGLIKE_ATOL


def test_hminus_limits(case_factory):
    """The H- continuum (rows per wavenumber, weights per level) enters the
    single-point, batched and adjoint kernels (parity in
    test_emission_parity, test_fp64_batch_with_hminus_and_clouds and the
    gradient tests); the adjoint that recomputes its forward sweep (when the
    kept quantities exceed the adjoint memory) does not take it and refuses
    the case with InvalidInput."""
    r, o, params, _ = retrieval(case_factory, "hminus")
    pts = np.array(params[:2])
    r.set_adjoint_memory(0)                    # no kept quantities: the recomputing adjoint
    try:
        with pytest.raises(api.BearError) as e:
            r.gradientBatch(pts)
        assert e.value.code == -1 and "H-" in str(e.value)
    finally:
        r.set_adjoint_memory(96 << 30)
    assert abs(r.computeLikelihood(params[0]) - o.evaluate(params[0])["loglike"]) <= LOGLIKE_ATOL


def test_cloud_limits(case_factory):
    """Cloud layers (a grey and a power-law model) enter the single-point,
    batched and emission adjoint kernels (parity in test_emission_parity,
    test_fp64_batch_with_hminus_and_clouds and the gradient tests); the
    recomputing adjoint and the transmission adjoint with clouds are refused
    with InvalidInput.  A cloud with zero
    optical depth leaves the spectrum as without it."""
    r, o, params, _ = retrieval(case_factory, "cloudy_tiny")
    pts = np.array(params[:2])
    r.set_adjoint_memory(0)                    # no kept quantities: the recomputing adjoint
    try:
        with pytest.raises(api.BearError) as e:
            r.gradientBatch(pts)
        assert e.value.code == -1 and "cloud" in str(e.value)
    finally:
        r.set_adjoint_memory(96 << 30)
    rk, _, pk, _ = retrieval(case_factory, "transmission_cloudy")
    with pytest.raises(api.BearError) as e:
        rk.gradientBatch(np.array(pk[:1]))
    assert e.value.code == -1 and "cloud" in str(e.value)
    p = np.array(params[0])
    clear = p.copy()
    clear[3 + 2 + 7] = 0.0                     # grey tau
    clear[3 + 2 + 7 + 3] = 0.0                 # power-law tau
    ref = o.evaluate(clear)
    out = r.computeModel(clear, True)
    rel = np.max(np.abs(out.spectrum - ref["spectrum"]) / np.abs(ref["spectrum"]))
    assert rel <= SPECTRUM_RTOL
    cloudy = r.computeModel(p, True)
    assert np.max(np.abs(cloudy.spectrum / out.spectrum - 1)) > 1e-3   # the clouds matter
