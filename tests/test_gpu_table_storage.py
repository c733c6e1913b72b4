"""Tables stored at reduced precision only (bear_global_config.table_storage):
each species' FP64 rows are split into the per-wavenumber FP32 base and the
residuals while the tables load, and no FP64 rows stay on the device.  The
batched results equal, bit for bit, those of FP64 storage split afterwards
by bear_retrieval_set_table_precision (the same split of the same values);
single-point likelihoods and models run as batches of one; the calls that
need FP64 rows are refused; the device holds the residuals and the base
only."""

import numpy as np
import pytest

from bear_amd import api, synth

pytestmark = pytest.mark.gpu

CASES = ("emission_c1", "transmission_plain", "eclipse_small")
BATCH = 5

_OPEN = {}


def opened(case_factory, name, storage):
    key = (name, storage)
    if key not in _OPEN:
        folder, params = case_factory(name)
        case = synth.CASES[name]()
        pts = np.array([params[0]] + [synth.perturbed(params[0], case, seed=300 + k, scale=2.0)
                                      for k in range(BATCH - 1)])
        _OPEN[key] = (synth.open_retrieval(folder, table_storage=storage), pts)
    return _OPEN[key]


@pytest.mark.parametrize("precision", ["fp32", "fp16", "bf16", "int16"])
@pytest.mark.parametrize("name", CASES)
def test_reduced_storage_equals_split_after_loading(case_factory, name, precision):
    full, pts = opened(case_factory, name, "fp64")
    low, _ = opened(case_factory, name, precision)
    print(f"[gpu] table storage {name} {precision}", flush=True)
    full.set_table_precision(precision)
    try:
        spec_a, bins_a, neg_a = full.computeModelBatch(pts)
        ll_a = full.computeLikelihoodBatch(pts)
    finally:
        full.set_table_precision("fp64")
    spec_b, bins_b, neg_b = low.computeModelBatch(pts)
    ll_b = low.computeLikelihoodBatch(pts)
    np.testing.assert_array_equal(spec_b, spec_a)
    for a, b in zip(bins_a, bins_b):
        np.testing.assert_array_equal(b, a)
    np.testing.assert_array_equal(neg_b, neg_a)
    np.testing.assert_array_equal(ll_b, ll_a)
    # the single-point calls: batches of one
    for k in (0, 3):
        assert low.computeLikelihood(pts[k]) == ll_b[k]
        out = low.computeModel(pts[k])
        np.testing.assert_array_equal(out.spectrum, spec_b[k])
        for j, b in enumerate(bins_b):
            np.testing.assert_array_equal(out.spectrum_obs[j], b[k])
    # against the FP64 single-point model, in the same units
    ref = full.computeModel(pts[0])
    rel = np.max(np.abs(spec_b[0] - ref.spectrum) / np.abs(ref.spectrum))
    print(f"[gpu] {name} {precision} storage: spectrum rel {rel:.2e} against FP64", flush=True)
    assert rel < {"fp32": 1e-5, "fp16": 3e-3, "bf16": 2e-2, "int16": 2e-4}[precision]


@pytest.mark.parametrize("precision", ["fp16", "bf16", "int16"])
def test_reduced_storage_device_memory(case_factory, precision):
    full, pts = opened(case_factory, "emission_c1", "fp64")
    low, _ = opened(case_factory, "emission_c1", precision)
    low.computeLikelihood(pts[0])
    full.computeLikelihood(pts[0])
    fb, lb = full.stats()["table_device_bytes"], low.stats()["table_device_bytes"]
    n = low.stats()["nb_spectral_points"]
    print(f"[gpu] table bytes: fp64 {fb}, {precision} only {lb}", flush=True)
    # 2 bytes per value plus per species slot an FP32 base row (int16: and
    # a scale row), at most 16 slots, rows padded by less than 1024 values
    assert fb // 4 < lb <= fb // 4 + 2 * 16 * 4 * (n + 1024)
    assert lb < fb / 3.5


def test_reduced_storage_refuses_fp64_calls(case_factory):
    import torch
    low, pts = opened(case_factory, "emission_c1", "fp16")
    for p in ("fp64", "fp32", "bf16"):
        with pytest.raises(api.BearError) as e:
            low.set_table_precision(p)
        assert e.value.code == -1 and "reduced precision only" in str(e.value)
    low.set_table_precision("fp16")                    # its own precision: nothing to do
    with pytest.raises(api.BearError) as e:
        low.gradientBatch(pts[:2])
    assert e.value.code == -1
    spec = torch.zeros(low.nb_spectral_points, dtype=torch.float64, device="cuda")
    for no_alloc in (False, True):
        with pytest.raises(api.BearError) as e:
            low.calcModelGPU(pts[0], spec.data_ptr(), [], no_alloc=no_alloc)
        assert e.value.code == -1 and "FP64" in str(e.value)
    # still usable after the refusals
    assert np.isfinite(low.computeLikelihood(pts[1]))
