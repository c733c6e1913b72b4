"""Full-size parity of the HIP engine with the CPU oracle on the
configurations' own workloads, through the C ABI:

  emission_c2      1 evaluation, 100 levels, 6 absorbers + 2 CIA + H2
                   Rayleigh, 400,000 wavenumbers, 300 bandpass bins
  transmission_c3  256 live points, 200 levels, 8 absorbers + 2 CIA + H2
                   Rayleigh, 300,000 wavenumbers, 20 / 147 / 400 bins
  emission_c4      64 live points, the emission_c2 model at 200,000
                   wavenumbers, 16 parameters (value and gradient)
  eclipse_c5       512 live points, 100 levels, 12 absorbers + 2 CIA on
                   30 T x 12 p tables, 1,000,000 wavenumbers, 9 bandpasses
                   + a 400-bin spectrum; FP16 and BF16 tables

Every wavenumber of each configuration's list is sampled (the observations
tile the list).  Tables are generated on the GPU and held in memory by both
the engine and the oracle.  Bars: spectra and bins within 1e-5 relative,
log-likelihood within 1e-6 absolute of the oracle (FP64 tables); reduced
precision: at most 10 ppm of transit or eclipse depth against the FP64
oracle; gradients within 1e-4 of the oracle's central differences.  Also
the size-independent closed forms at full size."""

import math
import os
import time

import numpy as np
import pytest

from bear_amd import synth
from oracle import api as oracle_api

pytestmark = pytest.mark.gpu

SPECTRUM_RTOL = 1e-5
LOGLIKE_ATOL = 1e-6
PPM_BAR = 10.0
GRADIENT_RTOL = 1e-4

# name: (builder, live points, full wavenumber count)
CONFIGS = {
    "emission_c2": ("emission_c2", 16, 400_000),
    "transmission_c3": ("transmission_c3", 256, 300_000),
    "emission_c4": ("emission_c4", 64, 200_000),
    "eclipse_c5": ("eclipse_c5", 512, 1_000_000),
}


def log(msg):
    print(f"[gpu] {msg}", flush=True)


class Full:
    """One configuration at full size: the case (its tables stay alive for
    the oracle, which samples rows on first use), the oracle, the engine and
    the live points.  The observed data are the oracle's model at the case's
    parameters plus 2 % noise."""

    def __init__(self, name, scratch):
        builder, self.batch, self.n = CONFIGS[name]
        self.name = name
        t0 = time.time()
        self.case = getattr(synth, builder)(backend="torch")
        log(f"{name}: tables generated in {time.time() - t0:.1f} s")
        self.folder = os.path.join(scratch, name)
        self.oracle = self._oracle()
        truth = self.case.parameters.copy()
        ref = self.oracle.evaluate(truth)
        synth.inject_data(self.case, self.input_order(ref["bins"]))
        self.oracle = self._oracle()
        t0 = time.time()
        self.engine = synth.open_retrieval_in_memory(self.case, os.path.join(scratch, name + "_engine"))
        log(f"{name}: engine set up in {time.time() - t0:.1f} s, {self.engine.nb_spectral_points} points")
        self.points = np.array([truth] + [synth.perturbed(truth, self.case, seed=2000 + k)
                                          for k in range(self.batch - 1)])
        self._refs = {}

    def _oracle(self):
        folder = synth.write_case(self.case, self.folder, tables=False)
        oracle_api.clear_tables()
        for s in self.case.species:
            oracle_api.register_table(s.folder, np.repeat(s.temperatures, s.pressures.size),
                                      np.tile(s.pressures, s.temperatures.size),
                                      s.values.reshape(-1, s.values.shape[-1]), s.mass, s.log_storage)
        o = oracle_api.EmissionOracle(folder)
        oracle_api.clear_tables()              # the oracle keeps its own references to the rows
        return o

    def input_order(self, bins):
        out = []
        for i, b in enumerate(bins):
            out.append(b[::-1].copy() if self.oracle.observation(i)["ascending"] else b)
        return out

    def reference(self, k):
        """Oracle result of live point k: spectrum, bins in input order, log L."""
        if k not in self._refs:
            t0 = time.time()
            r = self.oracle.evaluate(self.points[k])
            log(f"{self.name}: oracle point {k} in {time.time() - t0:.1f} s")
            self._refs[k] = (r["spectrum"], self.input_order(r["bins"]), r["loglike"], r["neglect"])
        return self._refs[k]

    def close(self):
        self.engine.close()
        self.engine = None
        self.oracle = None


_FULL = {}


@pytest.fixture(scope="module")
def full_factory(tmp_path_factory):
    scratch = str(tmp_path_factory.mktemp("full_size"))

    def get(name):
        # every configuration stays set up until the module ends (about 67 GB
        # of device and 45 GB of host memory at once): the parametrized tests
        # visit the configurations in turn
        if name not in _FULL:
            _FULL[name] = Full(name, scratch)
        return _FULL[name]
    yield get
    for f in _FULL.values():
        f.close()
    _FULL.clear()


def _compare(name, spec, bins, ll, ref, spectrum_rtol=SPECTRUM_RTOL, loglike_atol=LOGLIKE_ATOL):
    rspec, rbins, rll, _ = ref
    rel = float(np.max(np.abs(spec - rspec) / np.abs(rspec)))
    brel = max(float(np.max(np.abs(b - e) / np.abs(e))) for b, e in zip(bins, rbins))
    log(f"{name}: spectrum rel {rel:.2e}, bins rel {brel:.2e}, |dlogL| {abs(ll - rll):.2e}")
    assert rel <= spectrum_rtol
    assert brel <= spectrum_rtol
    assert abs(ll - rll) <= loglike_atol


# ---------------------------------------------------------------- parity, FP64
@pytest.mark.parametrize("name", list(CONFIGS))
def test_full_size_single_point_parity(full_factory, name):
    """The whole grid is sampled; two live points (the case's parameters and
    a nearby one) match the oracle."""
    f = full_factory(name)
    r, o = f.engine, f.oracle
    assert r.nb_spectral_points == f.n == o.nb_points
    nu, _ = r.spectral_grid()
    np.testing.assert_array_equal(nu, o.grid()[0])
    for k in (0, 1):
        out = r.computeModel(f.points[k], True)
        ll = r.computeLikelihood(f.points[k])
        ref = f.reference(k)
        assert out.neglect_model == ref[3]
        _compare(f"{name} point {k}", out.spectrum, out.spectrum_obs, ll, ref)


@pytest.mark.parametrize("name", list(CONFIGS))
def test_full_size_fp64_batch(full_factory, name):
    """The configuration's whole batch in one launch sequence (FP64 tables):
    every point's log-likelihood equals its single-point call, and the first
    two match the oracle."""
    f = full_factory(name)
    r = f.engine
    r.set_table_precision("fp64")
    t0 = time.time()
    ll = r.computeLikelihoodBatch(f.points)
    log(f"{name}: batch of {len(f.points)} (first call, tuning) in {time.time() - t0:.2f} s, "
        f"launch {r.batch_stats()}")
    single = np.array([r.computeLikelihood(p) for p in f.points])
    np.testing.assert_allclose(ll, single, rtol=1e-9, atol=1e-9)
    spec, bins, neglect = r.computeModelBatch(f.points[:2], True)
    for k in (0, 1):
        _compare(f"{name} batch point {k}", spec[k], [b[k] for b in bins], ll[k], f.reference(k))


def _isothermal(f, p, temperature):
    g0 = synth.nb_general_parameters(f.case)
    nb_mix = sum(len(c[1]) for c in f.case.chemistry if c[0] == "iso")
    q = p.copy()
    q[g0 + nb_mix] = temperature
    q[g0 + nb_mix + 1:g0 + nb_mix + 7] = 1.0
    return q


@pytest.mark.parametrize("name", list(CONFIGS))
def test_full_size_closed_forms(full_factory, name):
    """Emission: an isothermal atmosphere emits pi B(T) 1e-3 (R_J/d)^2 f at
    every wavenumber.  Eclipse: a planet isothermal at the star's temperature
    gives (Rp/Rs)^2 1e6 at every wavenumber and in every bin.  Transmission:
    every depth lies between the bare disc (R_p/R_*)^2 1e6 and the disc out
    to the top level, ((R_p + z_top)/R_*)^2 1e6."""
    f = full_factory(name)
    r = f.engine
    p = f.points[1]
    if f.case.model == "emission":
        q = _isothermal(f, p, 1300.0)
        out = r.computeModel(q, True)
        nu, lam = r.spectral_grid()
        h, c, kb = synth.PLANCK_H, synth.LIGHT_C, synth.BOLTZMANN_K
        b = 2 * h * c * c * nu ** 3 / np.expm1(h * c * nu / (kb * 1300.0))
        expected = math.pi * b * 1e-3 * (synth.RADIUS_JUPITER / q[2]) ** 2 * q[1] / lam ** 2 * 1e4
        err = float(np.max(np.abs(out.spectrum - expected) / expected))
        log(f"{name}: isothermal identity max rel err {err:.2e}")
        assert err <= 1e-10
    elif f.case.model == "secondary_eclipse":
        q = _isothermal(f, p, p[2])
        out = r.computeModel(q, True)
        expected = q[1] ** 2 * 1e6
        err = float(np.max(np.abs(out.spectrum - expected)) / expected)
        berr = max(float(np.max(np.abs(b - expected))) / expected for b in out.spectrum_obs)
        log(f"{name}: eclipse at T_star max rel err {err:.2e}, bins {berr:.2e}")
        assert err <= 1e-10 and berr <= 1e-10
    else:
        out = r.computeModel(p, True)
        f.oracle.evaluate(p)
        z_top = f.oracle.atmosphere()["altitude"][-1]
        lo = (p[1] / p[2]) ** 2 * 1e6
        hi = ((p[1] + z_top) / p[2]) ** 2 * 1e6
        log(f"{name}: depth in [{out.spectrum.min():.3f}, {out.spectrum.max():.3f}] ppm, bounds [{lo:.3f}, {hi:.3f}]")
        assert np.all(out.spectrum >= lo * (1 - 1e-12)) and np.all(out.spectrum <= hi * (1 + 1e-12))


# ---------------------------------------------------------------- reduced precision
REDUCED = [("emission_c2", "fp32"), ("transmission_c3", "fp32"), ("transmission_c3", "fp16"),
           ("transmission_c3", "bf16"), ("transmission_c3", "int16"), ("eclipse_c5", "fp16"),
           ("eclipse_c5", "bf16"), ("eclipse_c5", "int16")]


def _reduced(f, precision):
    r = f.engine
    r.set_table_precision(precision)
    try:
        ll = r.computeLikelihoodBatch(f.points[:2])
        spec, bins, _ = r.computeModelBatch(f.points[:2], True)
    finally:
        r.set_table_precision("fp64")
    return ll, spec, bins


@pytest.mark.parametrize("name,precision", REDUCED)
def test_full_size_reduced_precision_spectral_bar(full_factory, name, precision):
    """Batched evaluation from FP32 / FP16 / BF16 / scaled INT16 tables against the FP64
    oracle.  Transit and eclipse depths: at most 10 ppm of depth at every
    wavenumber and in every bin.  Emission (FP32 tables): the spectra bar,
    1e-5 relative."""
    f = full_factory(name)
    ll, spec, bins = _reduced(f, precision)
    for k in (0, 1):
        rspec, rbins, rll, _ = f.reference(k)
        dll = abs(ll[k] - rll)
        if f.case.model == "emission":
            rel = float(np.max(np.abs(spec[k] - rspec) / rspec))
            brel = max(float(np.max(np.abs(b[k] - e) / e)) for b, e in zip(bins, rbins))
            log(f"{name} {precision} point {k}: spectrum rel {rel:.2e}, bins rel {brel:.2e}, |dlogL| {dll:.2e}")
            assert rel <= SPECTRUM_RTOL and brel <= SPECTRUM_RTOL
        else:
            ppm = float(np.max(np.abs(spec[k] - rspec)))
            bppm = max(float(np.max(np.abs(b[k] - e))) for b, e in zip(bins, rbins))
            log(f"{name} {precision} point {k}: {ppm:.3e} ppm of depth (bins {bppm:.3e}), |dlogL| {dll:.2e}")
            assert ppm <= PPM_BAR and bppm <= PPM_BAR


@pytest.mark.xfail(reason="reduced-precision tables do not reach the 1e-6 log-likelihood bar: with 2 % errors "
                          "the log-likelihood needs bins to about 1e-9 relative, which only FP64 tables give "
                          "(the measured differences are printed)", strict=False)
@pytest.mark.parametrize("name,precision", REDUCED)
def test_full_size_reduced_precision_loglike_bar(full_factory, name, precision):
    f = full_factory(name)
    ll, _, _ = _reduced(f, precision)
    dll = max(abs(ll[k] - f.reference(k)[2]) for k in (0, 1))
    log(f"{name} {precision}: |dlogL| {dll:.3e} against the 1e-6 bar")
    assert dll <= LOGLIKE_ATOL


# ---------------------------------------------------------------- gradients
def _clear_of_table_nodes(f, p, margin):
    """A live point near p whose level temperatures all stay `margin` (relative)
    away from the tables' temperature nodes, so that central differences do
    not straddle a kink of the bilinear interpolation."""
    nodes = np.unique(np.concatenate([s.temperatures for s in f.case.species]))
    g0 = synth.nb_general_parameters(f.case)
    nb_mix = sum(len(c[1]) for c in f.case.chemistry if c[0] == "iso")
    q = p.copy()
    for k in range(200):
        f.oracle.evaluate(q)
        t = f.oracle.atmosphere()["temperature"]
        if np.min(np.abs(t[:, None] - nodes[None, :]) / t[:, None]) > margin:
            return q
        q[g0 + nb_mix] *= 1.0 + 1e-3 * (k + 1)
    raise AssertionError("no live point clear of the table nodes")


def _central_differences(o, p, indices, h_rel, frozen=None):
    """Central differences at steps h and h/2 combined by Richardson
    extrapolation, (4 D(h/2) - D(h)) / 3: near the likelihood's maximum the
    curvature is orders of magnitude above the slope, and a parameter that
    enters nonlinearly (the distance, through (R/d)^2) leaves an O(h^2) error
    of 1e-3 in a plain central difference."""
    if frozen is not None:
        o.freeze_stops(frozen)
    try:
        out = {}
        for k in indices:
            h = h_rel * (abs(p[k]) if p[k] != 0 else 1.0)
            d = []
            for step in (h, h / 2):
                e = np.zeros_like(p)
                e[k] = step
                d.append((o.evaluate(p + e)["loglike"] - o.evaluate(p - e)["loglike"]) / (2 * step))
            out[k] = (4 * d[1] - d[0]) / 3
            log(f"central differences in parameter {k}: {d[0]:.10e} (h), {d[1]:.10e} (h/2), "
                f"extrapolated {out[k]:.10e}")
    finally:
        if frozen is not None:
            o.freeze_stops(None)
    return out


def _check_gradient(label, grad, fd, p):
    """Every component whose d logL / d ln p is at least 1e-3 of the largest
    within 1e-4 relative of the oracle's; every component within 1e-4 of the
    largest, in d logL / d ln p (a component thousands of times smaller
    than the largest is a difference of large cancelling contributions, and
    the finite differences resolve it no better than that)."""
    scale = max(abs(fd[k] * p[k]) for k in fd)
    for k in fd:
        rel = abs(grad[k] - fd[k]) / abs(fd[k])
        of_largest = abs(grad[k] - fd[k]) * abs(p[k]) / scale
        significant = abs(fd[k] * p[k]) >= 1e-3 * scale
        log(f"{label} d logL / d p[{k}]: adjoint {grad[k]:.10e}, oracle {fd[k]:.10e}, rel {rel:.1e}, "
            f"{of_largest:.1e} of the largest{'' if significant else ' (small component)'}")
        assert of_largest <= GRADIENT_RTOL, k
        if significant:
            assert rel <= GRADIENT_RTOL, k


def test_full_size_emission_gradient(full_factory):
    """emission_c4: the hand-written adjoint of one point of the 64-point
    batch against the oracle's extrapolated central differences (h = 1e-4
    relative) in every one of the 16 parameters (_check_gradient), with the
    forward quantities kept in FP32 (the default) and recomputed in FP64."""
    f = full_factory("emission_c4")
    r = f.engine
    r.set_table_precision("fp64")
    p = _clear_of_table_nodes(f, f.points[1], 3e-4)
    pts = f.points.copy()
    pts[1] = p
    ll, grad = r.gradientBatch(pts)
    try:
        r.set_adjoint_memory(0)
        ll_r, grad_r = r.gradientBatch(p[None, :])
    finally:
        r.set_adjoint_memory(96 << 30)
    assert abs(ll[1] - f.oracle.evaluate(p)["loglike"]) <= LOGLIKE_ATOL
    assert ll_r[0] == ll[1]
    t0 = time.time()
    fd = _central_differences(f.oracle, p, range(p.size), 1e-4)
    log(f"emission_c4: {4 * p.size} oracle evaluations in {time.time() - t0:.1f} s")
    _check_gradient("emission_c4 kept", grad[1], fd, p)
    _check_gradient("emission_c4 recomputed", grad_r[0], fd, p)


def test_full_size_transmission_gradient(full_factory):
    """transmission_c3: the adjoint of one live point against the oracle's
    central differences with every tangent path's stopping segment held at
    the point's own (the derivative the adjoint computes: the cut-off's
    jumps are not differentiated), in every parameter, at 1e-4 relative."""
    f = full_factory("transmission_c3")
    r = f.engine
    r.set_table_precision("fp64")
    p = _clear_of_table_nodes(f, f.points[1], 3e-4)
    ll, grad = r.gradientBatch(p[None, :])
    o = f.oracle
    o.record_stops(True)
    base = o.evaluate(p)
    stops = o.stops()
    o.record_stops(False)
    assert abs(ll[0] - base["loglike"]) <= LOGLIKE_ATOL
    t0 = time.time()
    fd = _central_differences(o, p, range(p.size), 1e-4, frozen=stops)
    log(f"transmission_c3: {4 * p.size} frozen-stop oracle evaluations in {time.time() - t0:.1f} s")
    _check_gradient("transmission_c3", grad[0], fd, p)
