"""The no-allocation variant of ForwardModel::calcModelGPU
(bear_forward_model_calc_model_gpu_noalloc): its per-evaluation device
buffers live in the library's static scratch array instead of the engine's
own.  Through the C ABI it must give exactly what the allocating entry point
gives, call after call, whatever the previous call on the same GPU left in
the array, and refuse a model whose buffers do not fit."""

import numpy as np
import pytest

from bear_amd import api, synth

pytestmark = pytest.mark.gpu


def _calc(r, p, no_alloc):
    import torch
    spec = torch.full((r.nb_spectral_points,), float("nan"), dtype=torch.float64, device="cuda")
    obs = [torch.full((n,), float("nan"), dtype=torch.float64, device="cuda") for n in r.nb_observation_points]
    torch.cuda.synchronize()
    neglect = r.calcModelGPU(p, spec.data_ptr(), [t.data_ptr() for t in obs], no_alloc=no_alloc)
    return neglect, spec.cpu().numpy(), [t.cpu().numpy() for t in obs]


def test_no_alloc_equals_calc_model_gpu(case_factory):
    """Emission, transmission, eclipse: each model's no-allocation call
    equals the allocating call bit for bit, the calls following one another
    from the largest grid to the smallest (each leaves a larger model's
    values in the array for the next), then the largest again."""
    names = ["emission_c1", "transmission_plain", "eclipse_small", "tiny", "emission_c1"]
    opened = {}
    for k, name in enumerate(names):
        if name not in opened:
            folder, params = case_factory(name)
            opened[name] = (synth.open_retrieval(folder), params)
        r, params = opened[name]
        p = params[1] if k < len(names) - 1 else params[2]
        print(f"[gpu] no-alloc call {k}: {name}, {r.nb_spectral_points} points", flush=True)
        neg_a, spec_a, obs_a = _calc(r, p, no_alloc=False)
        neg_b, spec_b, obs_b = _calc(r, p, no_alloc=True)
        assert neg_a == neg_b
        assert np.all(np.isfinite(spec_b))
        np.testing.assert_array_equal(spec_b, spec_a)
        for a, b in zip(obs_a, obs_b):
            np.testing.assert_array_equal(b, a)
        # the log-likelihood path still agrees afterwards
        assert r.computeLikelihood(p) == r.computeLikelihood(p)
    for r, _ in opened.values():
        r.close()


def test_no_alloc_refuses_a_model_beyond_the_scratch(tmp_path):
    """2.4 million spectral points need more than the 16 MiB array: the
    call is refused with InvalidInput, and the allocating call still works."""
    n = 2_400_000
    nu = 1000.0 + 0.01 * np.arange(n, dtype=np.float64)
    temps = np.array([500.0, 2500.0])
    press = np.array([1e-6, 1e2])
    values = np.full((2, 2, n), -25.0, dtype=np.float32)
    table = synth.SpeciesTable("H2O", "H2O", temps, press, values, log_storage=True)
    lam_lo, lam_hi = 1e4 / nu[-1], 1e4 / nu[0]
    edges = np.linspace(lam_hi * 0.999, lam_lo * 1.001, 11)
    be = np.stack([edges[:-1], edges[1:]], axis=1)
    obs = synth.ObservationSpec("wide", "band-spectroscopy", bin_edges=be, data=np.ones(10), error=np.full(10, 0.1))
    params = np.array([4.5, 1.0, 10 * synth.PARSEC, 1e-4, 1200.0])
    case = synth.Case("beyond_scratch", nu, 0.01, 10, 100.0, 1e-6, ("const",), [("iso", ["H2O"]), ("bg", ["H2He"])],
                      [table], [], [obs], params)
    r = synth.open_retrieval_in_memory(case, str(tmp_path / "beyond"))
    print(f"[gpu] no-alloc refusal: {r.nb_spectral_points} points", flush=True)
    assert r.nb_spectral_points > 2_100_000
    import torch
    spec = torch.zeros(r.nb_spectral_points, dtype=torch.float64, device="cuda")
    obs_dev = [torch.zeros(k, dtype=torch.float64, device="cuda") for k in r.nb_observation_points]
    torch.cuda.synchronize()
    with pytest.raises(api.BearError) as e:
        r.calcModelGPU(params, spec.data_ptr(), [t.data_ptr() for t in obs_dev], no_alloc=True)
    assert e.value.code == -1 and "16 MiB" in str(e.value)
    assert not r.calcModelGPU(params, spec.data_ptr(), [t.data_ptr() for t in obs_dev])
    assert np.all(np.isfinite(spec.cpu().numpy()))
    r.close()
