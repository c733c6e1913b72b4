"""The CPU oracle, pinned against the reference's own CPU code (golden
fixtures made by tests/golden/make_golden.py and, where the reference build
exists, a live comparison) and against closed forms."""

import json
import math
import os

import numpy as np
import pytest

from oracle import api
from bear_amd import synth

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _golden(name):
    fname = f"emission_{name}.npz" if name.startswith("tiny") else f"{name}.npz"
    return np.load(os.path.join(GOLDEN, fname))


ALL_CASES = ["tiny", "emission_c1", "eclipse_small", "transmission_plain", "transmission_variable_gravity",
             "transmission_fit_mmw", "transmission_fit_scale_height", "tiny_const_wavelength",
             "tiny_const_resolution", "many_species", "hminus", "transmission_hminus", "cloudy_tiny",
             "cloudy_eclipse", "transmission_cloudy"]


@pytest.mark.parametrize("name", ALL_CASES)
def test_oracle_matches_reference_fixtures(case_factory, name):
    from golden.make_golden import table_digest
    folder, params = case_factory(name)
    g = _golden(name)
    assert str(g["table_digest"]) == table_digest(folder), "regenerated tables differ from the fixture's"
    np.testing.assert_array_equal(np.array(params), g["parameters"])
    transmission = name.startswith("transmission")
    o = api.EmissionOracle(folder)
    nu, lam, idx = o.grid()
    np.testing.assert_array_equal(nu, g["wavenumber"])
    np.testing.assert_array_equal(idx, g["index_list"])
    for i in range(o.nb_observations):
        ob = o.observation(i)
        np.testing.assert_array_equal(ob["edge_indices"], g[f"obs{i}_edge_indices"].astype(np.int64))
        np.testing.assert_array_equal(ob["index_range"], g[f"obs{i}_index_range"].astype(np.int64))
        np.testing.assert_array_equal(ob["data"], g[f"obs{i}_data"])
        np.testing.assert_array_equal(ob["weight"], g[f"obs{i}_weight"])
        if f"obs{i}_conv_intervals" in g:
            np.testing.assert_array_equal(ob["conv"], g[f"obs{i}_conv_intervals"].astype(np.int64))
            np.testing.assert_array_equal(ob["sigma"], g[f"obs{i}_sigma"])
        if f"obs{i}_filter" in g:
            np.testing.assert_array_equal(ob["filter"], g[f"obs{i}_filter"])
            assert ob["filter_norm"] == g[f"obs{i}_filter_norm"][0]
    for k, p in enumerate(params):
        r = o.evaluate(p)
        pre = f"p{k}_"
        a = o.atmosphere()
        np.testing.assert_allclose(a["temperature"], g[pre + "temperature"], rtol=1e-14)
        np.testing.assert_allclose(a["altitude"], g[pre + "altitude"], rtol=1e-13)
        if pre + "cloud_optical_depth" in g:
            np.testing.assert_allclose(o.scale_height(), g[pre + "scale_height"], rtol=1e-13)
            np.testing.assert_allclose(o.cloud_optical_depths()[:, ::10], g[pre + "cloud_optical_depth"], rtol=1e-13)
            if transmission:
                np.testing.assert_allclose(o.cloud_extinction()[:, ::10], g[pre + "cloud_extinction"], rtol=1e-13)
        if transmission:
            if pre + "absorption_coeff" in g:
                kap, sca = o.coefficients()
                np.testing.assert_allclose(kap.T[::5], g[pre + "absorption_coeff"], rtol=1e-13)
                np.testing.assert_allclose(sca.T[::5], g[pre + "scattering_coeff"], rtol=1e-13)
            continue
        flux = r["spectrum_flux"] if k == 0 else r["spectrum_flux"][::10]
        np.testing.assert_allclose(flux, g[pre + "spectrum_flux"], rtol=1e-13, atol=0)
        np.testing.assert_allclose(r["spectrum"][::10], g[pre + "spectrum"], rtol=1e-13, atol=0)
        for i in range(o.nb_observations):
            np.testing.assert_allclose(r["bins"][i], g[pre + f"obs{i}_bins"], rtol=1e-13)
        ref_ll, ref_neglect = g[pre + "loglike"]
        assert r["neglect"] == bool(ref_neglect)
        assert abs(r["loglike"] - ref_ll) <= 1e-9 * max(1.0, abs(ref_ll))


@pytest.mark.skipif(not os.path.exists(api.REF_BIN), reason="reference build (oracle/ref.mk) not present")
@pytest.mark.parametrize("name", ["tiny", "eclipse_small", "transmission_plain", "transmission_variable_gravity",
                                  "tiny_const_wavelength", "tiny_const_resolution", "hminus", "many_species",
                                  "transmission_hminus", "cloudy_tiny", "cloudy_eclipse", "transmission_cloudy"])
def test_oracle_intermediates_match_live_reference(case_factory, tmp_path, name):
    folder, params = case_factory(name)
    ref = api.run_reference(folder, str(tmp_path / "ref"))
    o = api.EmissionOracle(folder)
    meta = json.load(open(os.path.join(folder, "case.json")))
    r = o.evaluate(np.array(meta["parameters"]))
    k, s = o.coefficients()
    np.testing.assert_array_equal(k.T, ref["absorption_coeff"])
    np.testing.assert_array_equal(s.T, ref["scattering_coeff"])
    np.testing.assert_array_equal(o.atmosphere()["number_densities"], ref["number_densities"])
    np.testing.assert_array_equal(o.atmosphere()["altitude"], ref["altitude"])
    if "cloud_optical_depth" in ref:
        np.testing.assert_array_equal(o.scale_height(), ref["scale_height"])
        np.testing.assert_array_equal(o.cloud_optical_depths(), ref["cloud_optical_depth"])
        if name.startswith("transmission"):
            np.testing.assert_array_equal(o.cloud_extinction(), ref["cloud_extinction"])
    if not name.startswith("transmission"):
        np.testing.assert_array_equal(r["spectrum"], ref["spectrum"])


def _planck(t, nu):
    h, c, k = synth.PLANCK_H, synth.LIGHT_C, synth.BOLTZMANN_K
    return 2 * h * c * c * nu ** 3 / np.expm1(h * c * nu / (k * t))


def test_isothermal_emission_is_scaled_planck(case_factory):
    """I = B is a fixed point of the short-characteristics recurrence, so an
    isothermal atmosphere emits pi B(T) 1e-3 (R_J/d)^2 f at every wavenumber,
    whatever the opacities."""
    folder, params = case_factory("tiny")
    o = api.EmissionOracle(folder)
    p = params[3]
    r = o.evaluate(p)
    nu, _, _ = o.grid()
    scale = (synth.RADIUS_JUPITER / p[2]) ** 2 * p[1]
    expected = math.pi * _planck(p[5], nu) * 1e-3 * scale
    np.testing.assert_allclose(r["spectrum_flux"], expected, rtol=1e-11)
    np.testing.assert_allclose(o.atmosphere()["temperature"], p[5], rtol=1e-14)


def test_cold_model_is_neglected(case_factory):
    folder, params = case_factory("tiny")
    r = api.EmissionOracle(folder).evaluate(params[4])
    assert r["neglect"] and r["loglike"] == -1e30


def test_gaussian_log_likelihood_closed_form(case_factory):
    folder, params = case_factory("tiny")
    o = api.EmissionOracle(folder)
    r = o.evaluate(params[1])
    ll = 0.0
    for i in range(o.nb_observations):
        ob = o.observation(i)
        m = r["bins"][i] + (params[1][-1] if i == 1 else 0.0)   # obs 1 carries the shift parameter
        s2 = ob["error"] ** 2
        ll += np.sum(ob["weight"] * (-0.5 * np.log(2 * np.pi * s2) - 0.5 * (ob["data"] - m) ** 2 / s2))
    assert abs(ll - r["loglike"]) < 1e-9 * abs(ll)


def test_hydrostatic_altitude_isothermal(case_factory):
    """Isothermal, constant composition: z(p) = R T / (mu g) ln(p0 / p); the
    trapezoid in 1/rho converges to it at second order in the log-p step."""
    folder, params = case_factory("emission_c1")
    o = api.EmissionOracle(folder)
    p = params[0].copy()
    p[3 + 2 + 1:] = 1.0
    o.evaluate(p)
    a = o.atmosphere()
    g = 10 ** p[0]
    mu = a["mean_molecular_weight"]
    assert np.ptp(mu) / mu[0] < 1e-12
    z = 8.314472e7 * a["temperature"][0] / (mu[0] * g) * np.log(a["pressure"][0] / a["pressure"])
    np.testing.assert_allclose(a["altitude"][1:], z[1:], rtol=2e-2)
    assert np.all(np.diff(a["altitude"]) > 0)


def test_affine_log_table_is_interpolated_exactly(tmp_path):
    """A table whose log10 sigma is affine in (T, log10 p) is reproduced
    exactly by the bilinear interpolation at any point inside the grid."""
    case = synth.tiny_case()
    nu = case.wavenumber_full
    t, pr = case.species[0].temperatures, case.species[0].pressures
    a0, at, ap = -24.0, 2.0e-4, 0.3
    lg = a0 + at * t[:, None, None] + ap * np.log10(pr)[None, :, None] + 0.0 * nu[None, None, :]
    for s in case.species[:2]:
        s.values = lg.astype(np.float32)
        s.log_storage = True
        s.mass = 0.0
    case.species = case.species[:2]
    case.rayleigh = []
    folder = synth.write_case(case, str(tmp_path / "affine"))
    o = api.EmissionOracle(folder)
    o.evaluate(case.parameters)
    atm = o.atmosphere()
    k, _ = o.coefficients()
    nd = atm["number_densities"]
    T, P = atm["temperature"], atm["pressure"]
    inside = (T > t[0]) & (T < t[-1]) & (P > pr[0]) & (P < pr[-1])
    assert inside.sum() >= 3
    lg32 = lambda tt, pp: np.float64(np.float32(a0 + at * tt + ap * np.log10(pp)))
    for l in np.nonzero(inside)[0]:
        # exact on the float32 nodes: interpolate the stored corner values
        it = np.searchsorted(t, T[l]) - 1
        ip = np.searchsorted(pr, P[l]) - 1
        c = [[lg32(t[it + a], pr[ip + b]) for b in (0, 1)] for a in (0, 1)]
        fp = (np.log10(P[l]) - np.log10(pr[ip])) / (np.log10(pr[ip + 1]) - np.log10(pr[ip]))
        ft = (T[l] - t[it]) / (t[it + 1] - t[it])
        lo = c[0][0] + (c[0][1] - c[0][0]) * fp
        hi = c[1][0] + (c[1][1] - c[1][0]) * fp
        sigma = 10 ** (lo + (hi - lo) * ft)
        expected = sigma * (nd[l, 11] + nd[l, 14])                      # H2O and CH4 share the table
        assert abs(k[l, 0] - expected) <= 1e-12 * expected
        # and the affine function itself, to float32 table precision
        assert abs(np.log10(k[l, 0] / (nd[l, 11] + nd[l, 14])) - (a0 + at * T[l] + ap * np.log10(P[l]))) < 1e-5


# ---------------------------------------------------------------- transmission closed forms
def _affine_transmission_case(tmp_path, log_sigma, mixing=1e-4, tag="t"):
    """Transmission case whose two absorbers share a table
    log10 sigma = log_sigma(T, p) (affine in log10 p, so interpolated
    exactly), isothermal, no CIA and no Rayleigh."""
    case = synth.transmission_small()
    t, pr = case.species[0].temperatures, case.species[0].pressures
    nu = case.wavenumber_full
    lg = np.asarray(log_sigma(t[:, None, None], pr[None, :, None]) + 0.0 * nu[None, None, :])
    for s in case.species[:2]:
        s.values = lg.astype(np.float32)
        s.log_storage, s.mass = True, 0.0
    case.species = case.species[:2]
    case.rayleigh = []
    folder = synth.write_case(case, str(tmp_path / tag))
    p = case.parameters.copy()
    p[3] = p[4] = mixing
    return folder, p


def test_transparent_atmosphere_depth(tmp_path):
    """No absorbers: every chord transmits, h = 0 and depth = (R_p / R_*)^2 1e6."""
    folder, p = _affine_transmission_case(tmp_path, lambda t, pr: -30.0 + 0 * t * pr, mixing=0.0)
    r = api.EmissionOracle(folder).evaluate(p)
    np.testing.assert_allclose(r["spectrum"], (p[1] / p[2]) ** 2 * 1e6, rtol=1e-14)


def test_opaque_atmosphere_depth(tmp_path):
    """Every chord below the top level opaque: the trapezoid of 2 r dr is
    exact, so (R + h)^2 = (R + z_{L-2})^2 + (R + z_{L-2}) dz_{L-2} (the top
    level transmits)."""
    folder, p = _affine_transmission_case(tmp_path, lambda t, pr: -8.0 + 0 * t * pr, mixing=1e-2)
    o = api.EmissionOracle(folder)
    r = o.evaluate(p)
    z = o.atmosphere()["altitude"]
    R = p[1]
    r2 = (R + z[-2]) ** 2 + (R + z[-2]) * (z[-1] - z[-2])
    np.testing.assert_allclose(r["spectrum"], r2 / p[2] ** 2 * 1e6, rtol=1e-12)


def test_constant_extinction_telescopes(tmp_path):
    """log10 sigma = a - log10 p with an isothermal atmosphere gives the same
    extinction k at every level, so the chord sum telescopes: tau_t =
    2 k sqrt((R + z_top)^2 - (R + z_t)^2) (no cutoff reached)."""
    a = -24.0
    folder, p = _affine_transmission_case(tmp_path, lambda t, pr: a - np.log10(pr) + 0 * t, mixing=1e-4)
    o = api.EmissionOracle(folder)
    r = o.evaluate(p)
    atm = o.atmosphere()
    z, T, P = atm["altitude"], atm["temperature"][0], atm["pressure"]
    kap, _ = o.coefficients()
    n_tot = P * 1e6 / synth.BOLTZMANN_K / T
    k_expected = 2 * 1e-4 * n_tot * 10 ** a / P             # two absorbers, sigma = 10^a / p
    np.testing.assert_allclose(kap[:, 0], k_expected, rtol=2e-6)   # float32 table nodes
    k = kap[:, 0].mean()
    R = p[1]
    b = R + z
    tau = 2 * k * np.sqrt((R + z[-1]) ** 2 - b ** 2)
    assert tau.max() < 10
    trans = np.exp(-tau)
    H = np.sum(2 * ((R + z[:-1]) * (1 - trans[:-1]) + (R + z[1:]) * (1 - trans[1:])) * np.diff(z) * 0.5)
    depth = (np.sqrt(H + R * R)) ** 2 / p[2] ** 2 * 1e6
    np.testing.assert_allclose(r["spectrum"], depth, rtol=1e-6)


# ---------------------------------------------------------------- the tau > 10 cut-off
def test_tangent_sum_stops_after_the_first_sum_above_ten():
    """transmission_spectrum.cpp:123-137 adds a segment, then stops once the
    sum exceeds 10.  Segments of exactly 5 put a partial sum at 10.0: the
    path goes on (10 is not above 10) and stops after the third segment at
    15; a stop tested before the add, or at >= 10, stops at 10."""
    tau, last = api.tangent_sum([1.0, 1.0, 1.0, 1.0], [2.5, 2.5, 2.5, 2.5, 2.5])
    assert (tau, last) == (15.0, 2)
    tau, last = api.tangent_sum([1.0, 1.0, 1.0], [1.0, 1.0, 1.0, 1.0])
    assert (tau, last) == (6.0, 2)                     # never above 10: every segment
    tau, last = api.tangent_sum([4.0, 1.0, 1.0], [0.25, 1.0, 1.0, 1.0])
    assert (tau, last) == (9.0, 2)
    tau, last = api.tangent_sum([4.0, 1.0, 1.0], [2.0, 1.0, 1.0, 1.0])
    assert (tau, last) == (12.0, 0)                    # the first segment alone passes 10
    tau, last = api.tangent_sum([1.0, 1.0, 1.0, 1.0], [2.5] * 5, frozen=1)
    assert (tau, last) == (10.0, 1)                    # a frozen stop overrides the rule


def test_constant_extinction_cutoff_closed_form(tmp_path):
    """log10 sigma = a - log10 p on exact float32 nodes and an isothermal
    atmosphere give the same extinction k = 2 x 1e6 10^a / (k_B T) at every
    level (two absorbers), so a path's partial sums telescope: after segment
    q, tau_t = 2 k sqrt((R + z_{q+1})^2 - (R + z_t)^2).  The path stops at
    the first q where that exceeds 10 and keeps that sum; the transit depth
    follows from the trapezoid of 2 r (1 - e^-tau) dr.  The case is chosen so
    that some paths stop early and others run to the top."""
    a, x = -24.0, 4e-4
    folder, p = _affine_transmission_case(tmp_path, lambda t, pr: a - np.log10(pr) + 0 * t, mixing=x, tag="cut")
    o = api.EmissionOracle(folder)
    r = o.evaluate(p)
    atm = o.atmosphere()
    T = atm["temperature"][0]
    k = 2 * x * 1e6 * 10 ** a / (synth.BOLTZMANN_K * T)
    kap, _ = o.coefficients()
    np.testing.assert_allclose(kap, k, rtol=1e-12)
    z, R = atm["altitude"], p[1]
    L = z.size
    tau, stop = np.empty(L - 1), np.empty(L - 1, int)
    margin = np.inf
    for t in range(L - 1):
        partial = 2 * k * np.sqrt((R + z[t + 1:]) ** 2 - (R + z[t]) ** 2)     # after segments t .. L-2
        above = np.nonzero(partial > 10)[0]
        q = above[0] if above.size else L - 2 - t
        tau[t], stop[t] = partial[q], t + q
        margin = min(margin, np.min(np.abs(partial / 10 - 1)))
    assert margin > 1e-6                                 # no partial sum within rounding of 10
    early = stop < L - 2
    assert 3 <= early.sum() <= L - 4, early.sum()        # both kinds of path
    for w in (0, o.nb_points // 2, o.nb_points - 1):
        t_or, s_or = o.paths(w)
        np.testing.assert_array_equal(s_or, stop)
        np.testing.assert_allclose(t_or, tau, rtol=1e-9)
    trans = np.append(np.exp(-tau), 1.0)
    H = np.sum(((R + z[:-1]) * (1 - trans[:-1]) + (R + z[1:]) * (1 - trans[1:])) * np.diff(z))
    np.testing.assert_allclose(r["spectrum"], (H + R * R) / p[2] ** 2 * 1e6, rtol=1e-10)


# ---------------------------------------------------------------- eclipse and likelihood closed forms
def test_isothermal_eclipse_at_the_star_temperature(case_factory):
    """A planet isothermal at the star's temperature emits pi B(T_*) 1e-3, the
    blackbody star's flux (star_blackbody.cpp:40-54), so every high-resolution
    point and every bin, whatever its filter, line-spread function or band,
    is (Rp/Rs)^2 1e6 (secondary_eclipse.cpp:231-251)."""
    folder, params = case_factory("eclipse_small")
    o = api.EmissionOracle(folder)
    p = params[0].copy()
    nb_mix = 2
    p[3 + nb_mix] = p[2]                  # T_0 = T_star
    p[3 + nb_mix + 1:3 + nb_mix + 7] = 1.0
    r = o.evaluate(p)
    np.testing.assert_allclose(o.atmosphere()["temperature"], p[2], rtol=1e-14)
    expected = p[1] ** 2 * 1e6
    np.testing.assert_allclose(r["spectrum"], expected, rtol=1e-10)
    for b in r["bins"]:
        np.testing.assert_allclose(b, expected, rtol=1e-10)
    # and half the star's radius ratio: a quarter of the depth
    q = p.copy()
    q[1] *= 0.5
    np.testing.assert_allclose(o.evaluate(q)["bins"][0], expected / 4, rtol=1e-10)


def test_gaussian_log_likelihood_with_error_inflation(case_factory, tmp_path):
    """With error inflation the last parameter e adds 10^e to every variance
    (retrieval.cpp:330-358): log L = sum_j w_j (-ln(2 pi s_j^2) / 2 -
    (d_j - m_j)^2 / (2 s_j^2)), s_j^2 = sigma_j^2 + 10^e."""
    import shutil
    src, params = case_factory("tiny")
    folder = str(tmp_path / "inflated") + "/"
    shutil.copytree(src, folder)
    lines = open(folder + "case.txt").read().splitlines()
    lines = [l if not l.startswith("error_inflation") else "error_inflation 1" for l in lines]
    open(folder + "case.txt", "w").write("\n".join(lines) + "\n")
    plain = api.EmissionOracle(src)
    o = api.EmissionOracle(folder)
    assert o.nb_parameters == plain.nb_parameters + 1
    for e in (-30.0, -20.0, -19.0):
        p = np.append(params[1], e)
        r = o.evaluate(p)
        ll = 0.0
        for i in range(o.nb_observations):
            ob = o.observation(i)
            m = r["bins"][i]
            s2 = ob["error"] ** 2 + 10.0 ** e
            ll += np.sum(ob["weight"] * (-0.5 * np.log(2 * np.pi * s2) - 0.5 * (ob["data"] - m) ** 2 / s2))
        assert abs(ll - r["loglike"]) < 1e-9 * abs(ll), (e, ll, r["loglike"])
    # a negligible inflation leaves the plain likelihood
    assert abs(o.evaluate(np.append(params[1], -60.0))["loglike"] - plain.evaluate(params[1])["loglike"]) < 1e-9


def test_grid_discretisations_differ(case_factory):
    """The three samplings of the same observations (spectral_grid.cpp:158-236):
    constant wavenumber steps of 1.5 cm^-1, constant wavelength steps of 2 nm,
    constant resolving power 1500; each keeps the band edges."""
    grids = {}
    for name in ("tiny", "tiny_const_wavelength", "tiny_const_resolution"):
        folder, _ = case_factory(name)
        grids[name] = api.EmissionOracle(folder).grid()[0]
    dnu = np.diff(grids["tiny"])
    dlam = -np.diff(1e4 / grids["tiny_const_wavelength"])
    res = (1e4 / grids["tiny_const_resolution"])[:-1] / -np.diff(1e4 / grids["tiny_const_resolution"])
    # each step takes the last full-grid point within the resolution (the
    # forced band edges add a few shorter steps)
    assert np.median(dnu) == pytest.approx(1.5, rel=0.02)
    assert np.percentile(dlam, 95) <= 2e-3 and np.median(dlam) > 1e-3
    assert np.percentile(res, 5) >= 1500 * (1 - 1e-9) and np.median(res) < 2 * 1500
    assert len({g.size for g in grids.values()}) == 3


# ---------------------------------------------------------------- direct models
def test_direct_models_closed_forms(tmp_path):
    """pyBeAR's direct calcSpectrum in the oracle: an isothermal atmosphere
    emits pi B(T) 1e-3 (R/d)^2; in pure N2 (no absorber) the transit depth
    is the bare disc; the mean molecular weight is the mixing-ratio average of the
    given species (FixedChemistry, fixed_chemistry.cpp:71-89)."""
    case = synth.tiny_case()
    folder = synth.write_direct_case(case, str(tmp_path / "direct"), (2.0, 4.5))
    o = api.EmissionOracle(folder)
    L = o.nb_levels
    rng = np.random.default_rng(2)
    species = ["H2O", "CH4", "H2", "He"]
    p, _, x = synth.direct_structure(rng, L, species)
    r, d = synth.RADIUS_JUPITER, 10 * synth.PARSEC
    spec = o.direct_spectrum(2e3, r, d, p, np.full(L, 1400.0), species, x)
    nu = o.grid()[0]
    assert nu.min() >= 1e4 / 4.5 * 0.999 and nu.max() <= 1e4 / 2.0 * 1.001
    expected = np.pi * _planck(1400.0, nu) * 1e-3 * r * r / d / d
    np.testing.assert_allclose(spec, expected, rtol=1e-11)
    weights = np.array([18.01528, 16.04246, 2.01588, 4.002602])
    np.testing.assert_allclose(o.atmosphere()["mean_molecular_weight"], x @ weights, rtol=1e-13)
    np.testing.assert_allclose(o.atmosphere()["pressure"], p, rtol=0)
    case.model = "transmission"
    folder = synth.write_direct_case(case, str(tmp_path / "direct_t"), (2.0, 4.5))
    o = api.EmissionOracle(folder)
    ratio = 0.1
    spec = o.direct_spectrum(2e3, r, ratio, p, np.full(L, 900.0), ["N2"], np.ones((L, 1)))   # no opacity
    np.testing.assert_allclose(spec, ratio ** 2 * 1e6, rtol=1e-13)


@pytest.mark.parametrize("model", ["emission", "transmission", "secondary_eclipse"])
def test_oracle_direct_matches_reference_fixtures(tmp_path, model):
    """The oracle's direct calcSpectrum against the reference build's
    composition of the same classes (setAtmosphericStructure with
    FixedChemistry, OpacityCalculation, ShortCharacteristics, the star from
    a tabulated spectrum): altitudes, number densities, coefficients and
    spectra."""
    import cases
    from golden.make_golden import table_digest
    g = np.load(os.path.join(GOLDEN, f"direct_{model}.npz"))
    folder, inputs = cases.direct_case(model, str(tmp_path / model))
    assert str(g["table_digest"]) == table_digest(folder)
    o = api.EmissionOracle(folder)
    np.testing.assert_array_equal(o.grid()[0], g["wavenumber"])
    for k, d in enumerate(inputs):
        pre = f"d{k}_"
        np.testing.assert_array_equal(d["x"], g[pre + "mixing_ratios"])
        spec = o.direct_spectrum(d["g"], d["a"], d["b"], d["p"], d["t"], d["species"], d["x"], d["variable_gravity"],
                                 d["cloud"])
        atm = o.atmosphere()
        np.testing.assert_allclose(atm["altitude"], g[pre + "altitude"], rtol=1e-13)
        np.testing.assert_allclose(atm["number_densities"], g[pre + "number_densities"], rtol=1e-14)
        kap, sca = o.coefficients()
        np.testing.assert_allclose(kap.T[::7], g[pre + "absorption_coeff"], rtol=1e-13)
        np.testing.assert_allclose(sca.T[::7], g[pre + "scattering_coeff"], rtol=1e-13)
        if pre + "spectrum" in g:
            np.testing.assert_allclose(spec, g[pre + "spectrum"], rtol=1e-12)
        if pre + "cloud_extinction" in g:
            np.testing.assert_allclose(o.cloud_extinction()[:, ::7], g[pre + "cloud_extinction"], rtol=1e-13)


@pytest.mark.skipif(not os.path.exists(api.REF_BIN), reason="reference build (oracle/ref.mk) not present")
@pytest.mark.parametrize("model", ["emission", "secondary_eclipse", "transmission"])
@pytest.mark.parametrize("which", [1, 2])
def test_oracle_direct_matches_live_reference(tmp_path, model, which):
    """Input 2 carries a caller-made cloud (the reference's FixedCloudModel)."""
    import cases
    folder, inputs = cases.direct_case(model, str(tmp_path / model))
    o = api.EmissionOracle(folder)
    d = inputs[which]
    cases.write_direct_input(folder, d)
    ref = api.run_reference(folder, str(tmp_path / "ref"))
    spec = o.direct_spectrum(d["g"], d["a"], d["b"], d["p"], d["t"], d["species"], d["x"], d["variable_gravity"],
                             d["cloud"])
    if model != "transmission":
        np.testing.assert_array_equal(spec, ref["spectrum"])
    np.testing.assert_array_equal(o.atmosphere()["altitude"], ref["altitude"])
    if d["cloud"] is not None and model == "transmission":
        np.testing.assert_array_equal(o.cloud_extinction(), ref["cloud_extinction"])
