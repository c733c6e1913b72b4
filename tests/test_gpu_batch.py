"""The batched path: many live points per launch sequence, through the C ABI.

FP64 tables: every point of a batch matches the oracle at the single-point
bars (spectra and bins within 1e-5 relative, log-likelihood within 1e-6
absolute) and the single-point call.  Reduced-precision tables: the error
against the FP64 oracle is measured, printed, and held under the stated
bounds.  Also: graph replays with new inputs, uneven batch sizes, neglected
points inside a batch, and launch-shape independence."""

import numpy as np
import pytest

from bear_amd import api, synth
from oracle import api as oracle_api

pytestmark = pytest.mark.gpu

SPECTRUM_RTOL = 1e-5
LOGLIKE_ATOL = 1e-6

# Error bounds of the reduced-precision tables against the FP64 oracle on
# these cases (about 3x the measured maxima): relative error of the
# high-resolution spectrum and of the bins, absolute error in ppm of depth
# (eclipse and transmission), absolute error of the log-likelihood.
# Measured on MI355X: fp32 4.3e-6 / 1.6e-7 / 1.2e-3 ppm / 6.3e-4;
# fp16 1.1e-3 / 7.0e-4 / 2.05 ppm / 0.50; bf16 7.5e-3 / 4.7e-3 / 13.7 ppm / 3.3;
# int16 4.0e-5 / 5.5e-7 / 0.014 ppm / 7.8e-4.
REDUCED_BOUNDS = {
    "fp32": dict(spectrum=1e-5, bins=1e-6, ppm=0.01, loglike=3e-3),
    "fp16": dict(spectrum=3e-3, bins=2e-3, ppm=6.0, loglike=2.0),
    "bf16": dict(spectrum=2e-2, bins=1.5e-2, ppm=40.0, loglike=10.0),
    "int16": dict(spectrum=1.2e-4, bins=1.5e-6, ppm=0.05, loglike=3e-3),
}

_CACHE = {}


def setup(case_factory, name, batch):
    key = (name, batch)
    if key not in _CACHE:
        folder, params = case_factory(name)
        case = synth.CASES[name]()
        pts = [params[0]] + [synth.perturbed(params[0], case, seed=100 + k, scale=2.0) for k in range(batch - 1)]
        _CACHE[key] = (folder, np.array(pts))
    return _CACHE[key]


_OPEN = {}


def opened(folder):
    if folder not in _OPEN:
        _OPEN[folder] = (synth.open_retrieval(folder), oracle_api.EmissionOracle(folder))
    return _OPEN[folder]


def oracle_bins(o, ref):
    out = []
    for i in range(o.nb_observations):
        ob = o.observation(i)
        out.append(ref["bins"][i][::-1] if ob["ascending"] else ref["bins"][i])
    return out


@pytest.mark.parametrize("name", ["tiny", "emission_c1", "eclipse_small", "transmission_plain",
                                  "transmission_fit_scale_height", "emission_wide", "eclipse_wide"])
def test_fp64_batch_matches_oracle_and_single_calls(case_factory, name):
    folder, pts = setup(case_factory, name, 5)
    r, o = opened(folder)
    print(f"[gpu] batch fp64 {name} ({len(pts)} points)", flush=True)
    r.set_table_precision("fp64")
    ll = r.computeLikelihoodBatch(pts)
    spec, bins, neglect = r.computeModelBatch(pts, True)
    for k, p in enumerate(pts):
        ref = o.evaluate(p)
        assert bool(neglect[k]) == ref["neglect"]
        rel = np.max(np.abs(spec[k] - ref["spectrum"]) / np.abs(ref["spectrum"]))
        assert rel <= SPECTRUM_RTOL, (k, rel)
        for i, expect in enumerate(oracle_bins(o, ref)):
            np.testing.assert_allclose(bins[i][k], expect, rtol=SPECTRUM_RTOL, atol=0)
        assert abs(ll[k] - ref["loglike"]) <= LOGLIKE_ATOL, (k, ll[k], ref["loglike"])
        single = r.computeLikelihood(p)
        assert abs(ll[k] - single) <= 1e-9 * max(1.0, abs(single)), (k, ll[k], single)


@pytest.mark.parametrize("name", ["hminus", "cloudy_tiny", "cloudy_eclipse", "transmission_hminus",
                                  "transmission_cloudy"])
def test_fp64_batch_with_hminus_and_clouds(case_factory, name):
    """The H- continuum and the cloud models in the batched kernels (FP64
    tables; each point's clouds travel with its record): a batch of the
    case's own parameter vectors (clouds at several heights and depths) and
    two perturbed ones matches the oracle at the single-point bars and the
    single-point calls; reduced-precision tables refuse these models."""
    folder, params = case_factory(name)
    case = synth.CASES[name]()
    pts = np.array(list(params) + [synth.perturbed(params[0], case, seed=300 + k, scale=2.0) for k in range(2)])
    r, o = opened(folder)
    print(f"[gpu] batch fp64 with extras {name} ({len(pts)} points)", flush=True)
    r.set_table_precision("fp64")
    ll = r.computeLikelihoodBatch(pts)
    spec, bins, neglect = r.computeModelBatch(pts, True)
    for k, p in enumerate(pts):
        ref = o.evaluate(p)
        assert bool(neglect[k]) == ref["neglect"]
        if ref["neglect"]:
            assert ll[k] == -1e30
            continue
        rel = np.max(np.abs(spec[k] - ref["spectrum"]) / np.abs(ref["spectrum"]))
        assert rel <= SPECTRUM_RTOL, (k, rel)
        for i, expect in enumerate(oracle_bins(o, ref)):
            np.testing.assert_allclose(bins[i][k], expect, rtol=SPECTRUM_RTOL, atol=0)
        assert abs(ll[k] - ref["loglike"]) <= LOGLIKE_ATOL, (k, ll[k], ref["loglike"])
        single = r.computeModel(p, True)
        np.testing.assert_allclose(spec[k], single.spectrum, rtol=1e-11, atol=0)
        one = r.computeLikelihood(p)
        assert abs(ll[k] - one) <= 1e-9 * max(1.0, abs(one)), (k, ll[k], one)
    r.set_table_precision("fp32")
    try:
        with pytest.raises(api.BearError) as e:
            r.computeLikelihoodBatch(pts[:2])
        assert "FP64 tables" in str(e.value)
    finally:
        r.set_table_precision("fp64")


def test_neglected_point_inside_a_batch(case_factory):
    folder, params = case_factory("tiny")
    r, o = opened(folder)
    r.set_table_precision("fp64")
    cold = params[4]                     # T < 50 K everywhere: neglected
    pts = np.array([params[0], cold, params[1]])
    ll = r.computeLikelihoodBatch(pts)
    assert ll[1] == -1e30
    assert abs(ll[0] - o.evaluate(params[0])["loglike"]) <= LOGLIKE_ATOL
    assert abs(ll[2] - o.evaluate(params[1])["loglike"]) <= LOGLIKE_ATOL


def test_graph_replays_follow_new_inputs(case_factory):
    """One captured graph per batch size, replayed with different points,
    with eager single-point work in between: every replay matches."""
    folder, pts = setup(case_factory, "emission_c1", 6)
    r, o = opened(folder)
    r.set_table_precision("fp64")
    r.set_batch_options(use_graph=True, autotune=True)
    expect = np.array([o.evaluate(p)["loglike"] for p in pts])
    first = r.computeLikelihoodBatch(pts)
    assert r.batch_stats()["used_graph"] == 1
    r.computeLikelihood(pts[0])
    second = r.computeLikelihoodBatch(pts[::-1].copy())
    third = r.computeLikelihoodBatch(pts)
    np.testing.assert_allclose(first, expect, rtol=0, atol=LOGLIKE_ATOL)
    np.testing.assert_allclose(second, expect[::-1], rtol=0, atol=LOGLIKE_ATOL)
    np.testing.assert_array_equal(first, third)
    r.set_batch_options(use_graph=False, autotune=True)
    eager = r.computeLikelihoodBatch(pts)
    assert r.batch_stats()["used_graph"] == 0
    np.testing.assert_array_equal(eager, first)
    r.set_batch_options(use_graph=True, autotune=True)


@pytest.mark.parametrize("name", ["emission_c1", "transmission_plain"])
def test_uneven_batch_sizes_and_launch_shapes(case_factory, name):
    folder, pts = setup(case_factory, name, 7)
    r, o = opened(folder)
    r.set_table_precision("fp64")
    expect = np.array([o.evaluate(p)["loglike"] for p in pts])
    r.set_batch_options(use_graph=True, autotune=False)
    fixed = np.concatenate([r.computeLikelihoodBatch(pts[:1]), r.computeLikelihoodBatch(pts[1:4]),
                            r.computeLikelihoodBatch(pts[4:])])
    r.set_batch_options(use_graph=True, autotune=True)
    tuned = r.computeLikelihoodBatch(pts)
    print(f"[gpu] {name} tuned launch {r.batch_stats()}", flush=True)
    np.testing.assert_allclose(fixed, expect, rtol=0, atol=LOGLIKE_ATOL)
    np.testing.assert_allclose(tuned, expect, rtol=0, atol=LOGLIKE_ATOL)


@pytest.mark.parametrize("precision", ["fp32", "fp16", "bf16", "int16"])
@pytest.mark.parametrize("name", ["emission_c1", "eclipse_small", "transmission_plain"])
def test_reduced_precision_error_against_fp64_oracle(case_factory, name, precision):
    folder, pts = setup(case_factory, name, 4)
    r, o = opened(folder)
    r.set_table_precision(precision)
    try:
        ll = r.computeLikelihoodBatch(pts)
        spec, bins, _ = r.computeModelBatch(pts, True)
    finally:
        r.set_table_precision("fp64")
    err_spec = err_bins = err_ll = 0.0
    err_ppm = 0.0
    for k, p in enumerate(pts):
        ref = o.evaluate(p)
        err_spec = max(err_spec, np.max(np.abs(spec[k] - ref["spectrum"]) / np.abs(ref["spectrum"])))
        err_ppm = max(err_ppm, np.max(np.abs(spec[k] - ref["spectrum"])))
        for i, expect in enumerate(oracle_bins(o, ref)):
            err_bins = max(err_bins, np.max(np.abs(bins[i][k] - expect) / np.abs(expect)))
        err_ll = max(err_ll, abs(ll[k] - ref["loglike"]))
    unit = " (abs, ppm of depth)" if name != "emission_c1" else ""
    print(f"[gpu] {name} {precision}: spectrum rel {err_spec:.3e}, abs {err_ppm:.3e}{unit}, "
          f"bins rel {err_bins:.3e}, |dlogL| {err_ll:.3e}", flush=True)
    b = REDUCED_BOUNDS[precision]
    assert err_spec <= b["spectrum"]
    assert err_bins <= b["bins"]
    if name != "emission_c1":
        assert err_ppm <= b["ppm"]
    assert err_ll <= b["loglike"]
