"""The hand-written adjoint of the emission, secondary-eclipse and
transmission likelihoods, through the C ABI and the torch.autograd.Function.

Bar: every gradient component within 1e-4 relative of central finite
differences (step 1e-4 relative) of the GPU FP64 likelihood, and of the FP64
oracle's gradient (central differences at steps h and h/2 combined by
Richardson extrapolation), each component on its own scale (the parameters'
units differ by tens of decades), plus the finite differences' own error
estimate |D(h) - D(h/2)|."""

import os

import numpy as np
import pytest

from bear_amd import api, synth
from oracle import api as oracle_api

pytestmark = pytest.mark.gpu

RTOL = 1e-4
STEP = 1e-4


def build_inflated_tiny(folder):
    """The tiny emission case with error inflation (last parameter log10 of
    the added variance), data injected from the oracle."""
    case = synth.tiny_case()
    case.use_error_inflation = True
    # inflation of a third of the median variance
    errs = np.concatenate([ob.error for ob in case.observations])
    truth = np.concatenate([case.parameters, [np.log10(0.3 * np.median(errs) ** 2)]])
    case.parameters = truth
    folder = synth.write_case(case, folder)
    o = oracle_api.EmissionOracle(folder)
    r = o.evaluate(truth)
    bins = [b[::-1].copy() if o.observation(i)["ascending"] else b for i, b in enumerate(r["bins"])]
    synth.inject_data(case, bins)
    folder = synth.write_case(case, folder)
    return folder, [truth, synth.perturbed(truth, case, seed=21)]


_CASES = {}


def gradient_case(case_factory, tmp_path_factory, name):
    if name not in _CASES:
        if name == "tiny_inflated":
            folder, params = build_inflated_tiny(str(tmp_path_factory.mktemp(name)))
        else:
            folder, params = case_factory(name)
            params = params[:2]
        _CASES[name] = (folder, params, synth.open_retrieval(folder))
    return _CASES[name]


def steps(p):
    # relative steps; an absolute one of 1e-6 for parameters at 0 (observation
    # shifts: the log-likelihood is not quadratic in them, and at 1e-4 the
    # differences' truncation error reaches 1e-5 relative where the model is
    # far from the data)
    return STEP * np.where(p == 0, 1e-2, np.abs(p))


def central(f, p, h):
    g = np.empty_like(p)
    for k in range(p.size):
        e = np.zeros_like(p)
        e[k] = h[k]
        g[k] = (f(p + e) - f(p - e)) / (2 * h[k])
    return g


def check(grad, ref, label, fd_error):
    tol = RTOL * np.abs(ref) + 10 * fd_error + 1e-300
    rel = np.abs(grad - ref) / np.maximum(np.abs(ref), 1e-300)
    print(f"[gpu] {label}: max relative difference {rel.max():.2e} (component {rel.argmax()})", flush=True)
    bad = np.abs(grad - ref) > tol
    for k in np.where(bad)[0]:
        print(f"[gpu] {label} component {k}: adjoint {grad[k]:.10e} reference {ref[k]:.10e}", flush=True)
    assert not bad.any(), label


TRANSMISSION = ["transmission_plain", "transmission_variable_gravity", "transmission_fit_mmw",
                "transmission_fit_scale_height"]


@pytest.mark.parametrize("name", ["tiny", "tiny_inflated", "emission_c1", "eclipse_small", "eclipse_wide", "hminus",
                                  "cloudy_tiny", "cloudy_eclipse"] + TRANSMISSION + ["transmission_hminus"])
def test_adjoint_matches_central_differences(case_factory, tmp_path_factory, name):
    folder, params, r = gradient_case(case_factory, tmp_path_factory, name)
    r.set_table_precision("fp64")
    pts = np.array(params)
    ll, grad = r.gradientBatch(pts)
    for k, p in enumerate(pts):
        print(f"[gpu] gradient {name} point {k}", flush=True)
        assert abs(ll[k] - r.computeLikelihood(p)) <= 1e-9 * max(1.0, abs(ll[k]))
        h = steps(p)
        fd = central(r.computeLikelihood, p, h)
        fd2 = central(r.computeLikelihood, p, h / 2)
        check(grad[k], fd2, f"{name} point {k} vs GPU central differences", np.abs(fd - fd2))


@pytest.mark.parametrize("name", ["tiny", "tiny_inflated", "eclipse_small", "eclipse_wide", "transmission_plain",
                                  "transmission_fit_scale_height", "hminus", "transmission_hminus", "cloudy_tiny",
                                  "cloudy_eclipse"])
def test_adjoint_matches_oracle_fp64_gradient(case_factory, tmp_path_factory, name):
    folder, params, r = gradient_case(case_factory, tmp_path_factory, name)
    r.set_table_precision("fp64")
    o = oracle_api.EmissionOracle(folder)
    f = lambda q: o.evaluate(q)["loglike"]
    p = np.array(params[0])
    _, grad = r.gradientBatch(p[None, :])
    h = steps(p)
    d1, d2 = central(f, p, h), central(f, p, h / 2)
    ref = (4 * d2 - d1) / 3
    check(grad[0], ref, f"{name} vs oracle Richardson gradient", np.abs(ref - d2))


def stored_bytes_per_point(r):
    """Device bytes one live point takes when its forward quantities are
    kept: 2 intensities less the Planck function and the opacity x levels x
    padded spectral points in FP32, the per-tile partial sums of every
    (level, slot) and the flux adjoint in FP64."""
    st = r.stats()
    n, L = st["nb_spectral_points"], st["nb_levels"]
    stride = (n + 63) // 64 * 64
    slots = (st["nb_species"] + 3) // 4 * 4 + 4
    tiles = (n + 63) // 64
    return 4 * L * 3 * stride + 8 * (L * slots * tiles + stride)


@pytest.mark.parametrize("name", ["tiny", "emission_c1", "eclipse_small"])
def test_kept_chunked_and_recomputed_forward_agree(case_factory, tmp_path_factory, name):
    """The backward pass over kept forward quantities (whole batch, and one
    point per chunk) and over a forward sweep recomputed inside the kernel."""
    folder, params, r = gradient_case(case_factory, tmp_path_factory, name)
    r.set_table_precision("fp64")
    pts = np.array(params)
    try:
        ll_s, g_s = r.gradientBatch(pts)
        assert r.batch_stats()["adjoint_chunk"] == len(pts)
        r.set_adjoint_memory(stored_bytes_per_point(r) * 3 // 2)
        ll_c, g_c = r.gradientBatch(pts)
        assert r.batch_stats()["adjoint_chunk"] == 1
        r.set_adjoint_memory(0)
        ll_r, g_r = r.gradientBatch(pts)
        assert r.batch_stats()["adjoint_chunk"] == 0
    finally:
        r.set_adjoint_memory(96 << 30)
    print(f"[gpu] {name}: chunked max |diff| {np.max(np.abs(g_c - g_s)):.3e}, "
          f"recomputed max rel diff {np.max(np.abs(g_r - g_s) / np.maximum(np.abs(g_s), 1e-300)):.3e}", flush=True)
    np.testing.assert_array_equal(ll_c, ll_s)
    np.testing.assert_array_equal(g_c, g_s)
    np.testing.assert_array_equal(ll_r, ll_s)
    # the kept path reads the forward's I - B and opacity in FP32 where the
    # fused kernel recomputes them in FP64: the two gradients agree far
    # inside the 1e-4 bar, not bit for bit
    np.testing.assert_allclose(g_r, g_s, rtol=1e-5, atol=1e-6 * np.max(np.abs(g_s)))


def test_autograd_function_backward(case_factory, tmp_path_factory):
    import torch
    from bear_amd.autograd import log_likelihood
    folder, params, r = gradient_case(case_factory, tmp_path_factory, "tiny")
    r.set_table_precision("fp64")
    pts = np.array(params)
    x = torch.tensor(pts, dtype=torch.float64, requires_grad=True)
    ll = log_likelihood(r, x)
    w = torch.tensor([0.7, -1.3], dtype=torch.float64)
    (ll * w).sum().backward()
    ll_ref, grad_ref = r.gradientBatch(pts)
    np.testing.assert_array_equal(ll.detach().numpy(), ll_ref)
    np.testing.assert_allclose(x.grad.numpy(), w.numpy()[:, None] * grad_ref, rtol=1e-15, atol=0)
    plain = log_likelihood(r, torch.tensor(pts, dtype=torch.float64))
    np.testing.assert_array_equal(plain.numpy(), r.computeLikelihoodBatch(pts))
    # lazy: the forward is the plain batch, the adjoint runs in backward
    y = torch.tensor(pts, dtype=torch.float64, requires_grad=True)
    ll_lazy = log_likelihood(r, y, gradient="lazy")
    np.testing.assert_array_equal(ll_lazy.detach().numpy(), r.computeLikelihoodBatch(pts))
    (ll_lazy * w).sum().backward()
    np.testing.assert_allclose(y.grad.numpy(), w.numpy()[:, None] * grad_ref, rtol=1e-15, atol=0)
    z = torch.tensor(pts[0], dtype=torch.float64, requires_grad=True)
    log_likelihood(r, z).backward()
    np.testing.assert_allclose(z.grad.numpy(), grad_ref[0], rtol=1e-12, atol=0)


def test_gradient_refuses_other_models_and_precisions(case_factory, tmp_path_factory):
    folder, params, r = gradient_case(case_factory, tmp_path_factory, "tiny")
    r.set_table_precision("fp32")
    with pytest.raises(api.BearError) as e:
        r.gradientBatch(np.array(params))
    assert e.value.code == -1
    r.set_table_precision("fp64")
    # the transmission adjoint has no recomputing variant
    folder_t, params_t, rt = gradient_case(case_factory, tmp_path_factory, "transmission_plain")
    rt.set_adjoint_memory(0)
    try:
        with pytest.raises(api.BearError) as e:
            rt.gradientBatch(np.array(params_t[:1]))
        assert e.value.code == -1
    finally:
        rt.set_adjoint_memory(96 << 30)


def test_transmission_whole_and_chunked_batches_agree(case_factory, tmp_path_factory):
    """One chunk for the batch and one point per chunk: the same gradients."""
    folder, params, r = gradient_case(case_factory, tmp_path_factory, "transmission_plain")
    r.set_table_precision("fp64")
    pts = np.array(params)
    st = r.stats()
    n, L = st["nb_spectral_points"], st["nb_levels"]
    stride = (n + 63) // 64 * 64
    slots = (st["nb_species"] + 3) // 4 * 4 + 4
    rows = (L - 1 + 31) // 32 * 32
    per_point = 4 * L * 3 * stride + 8 * (L * slots * ((n + 63) // 64) + stride + 17 * (L - 1) ** 2 + (L - 1) * rows)
    try:
        ll_s, g_s = r.gradientBatch(pts)
        assert r.batch_stats()["adjoint_chunk"] == len(pts)
        r.set_adjoint_memory(per_point * 3 // 2)
        ll_c, g_c = r.gradientBatch(pts)
        assert r.batch_stats()["adjoint_chunk"] == 1
    finally:
        r.set_adjoint_memory(96 << 30)
    np.testing.assert_array_equal(ll_c, ll_s)
    np.testing.assert_array_equal(g_c, g_s)


def test_transmission_windows_match_the_whole_grid(case_factory, tmp_path_factory):
    """An adjoint memory below one point's kept forward quantities: one
    point at a time, its grid in windows of three chord ranges, the forward
    kernel rerun per window.  The same sums in the same order: the gradients
    equal the whole-grid path's bit for bit; below one range, an error."""
    folder, params, r = gradient_case(case_factory, tmp_path_factory, "transmission_plain")
    r.set_table_precision("fp64")
    pts = np.array(params)
    st = r.stats()
    n, L = st["nb_spectral_points"], st["nb_levels"]
    stride = (n + 63) // 64 * 64
    tiles = (n + 63) // 64
    slots = (st["nb_species"] + 3) // 4 * 4 + 4
    P = L - 1
    rows = (P + 31) // 32 * 32
    ranges = max(1, min(16, (n + 511) // 512))
    while True:                                          # as the engine: no empty range, multiples of 128
        npr = -(-(-(-n // ranges)) // 128) * 128
        if ranges == 1 or (ranges - 1) * npr < n:
            break
        ranges -= 1
    assert ranges >= 4
    fixed = 8 * (L * slots * tiles + stride + 17 * P * P + P * rows)
    kept_range = 4 * L * 3 * npr
    try:
        ll_s, g_s = r.gradientBatch(pts)
        assert r.batch_stats()["adjoint_chunk"] == len(pts)
        r.set_adjoint_memory(fixed + 3 * kept_range + kept_range // 2)
        print(f"[gpu] transmission windows: {ranges} ranges of {npr} wavenumbers, 3 per window", flush=True)
        ll_w, g_w = r.gradientBatch(pts)
        assert r.batch_stats()["adjoint_chunk"] == 1
        np.testing.assert_array_equal(ll_w, ll_s)
        np.testing.assert_array_equal(g_w, g_s)
        r.set_adjoint_memory(fixed + kept_range // 2)
        with pytest.raises(api.BearError) as e:
            r.gradientBatch(pts)
        assert e.value.code == -1 and "adjoint memory" in str(e.value)
    finally:
        r.set_adjoint_memory(96 << 30)


def test_neglected_point_has_zero_gradient(case_factory, tmp_path_factory):
    folder, params = case_factory("tiny")
    r = gradient_case(case_factory, tmp_path_factory, "tiny")[2]
    r.set_table_precision("fp64")
    cold = params[4]
    ll, grad = r.gradientBatch(np.array([params[0], cold]))
    assert ll[1] == -1e30
    assert np.all(grad[1] == 0)
    assert np.any(grad[0] != 0)
