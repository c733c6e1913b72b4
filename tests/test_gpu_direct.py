"""pyBeAR's direct forward models through the C ABI: EmissionModel,
TransmissionModel and OccultationModel on a grid sampled between two
wavelengths, whose calcSpectrum takes the atmosphere's structure (pressure,
temperature, mixing ratios per level) from the caller
(bear_python_wrapper.cpp:343-372, emission.cpp:268-340,
transmission.cpp:371-465, secondary_eclipse.cpp:395-482).  Parity with the
oracle's restatement within 1e-5 relative; closed forms on the GPU."""

import math

import numpy as np
import pytest

from bear_amd import api, synth
from oracle import api as oracle_api

pytestmark = pytest.mark.gpu

SPECTRUM_RTOL = 1e-5
RANGE = (2.0, 4.5)
LEVELS = 30
SPECIES = ["H2O", "CH4", "H2", "He"]
FASTCHEM = ["H2O1", "C1H4", "H2", "He"]
MODELS = ("emission", "transmission", "secondary_eclipse")
T_STAR = 5200.0


def _stellar(lam):
    """Blackbody star of T_STAR in W m^-2 micron^-1 at wavelengths lam (micron)."""
    return synth.blackbody_flux_lambda(lam, T_STAR, 1.0)


_OPEN = {}


def opened(tmp_path_factory, model):
    if model not in _OPEN:
        case = synth.tiny_case()
        case.model = model
        case.nb_levels = LEVELS
        stellar = None
        if model == "secondary_eclipse":
            lam = np.linspace(1.6, 5.1, 4001)
            stellar = (lam, _stellar(lam))
        folder = synth.write_direct_case(case, str(tmp_path_factory.mktemp("direct_" + model)), RANGE, stellar)
        cfg = api.GlobalConfig(retrieval_folder_path=folder, cross_section_file_path=folder + "opacity/",
                               wavenumber_file_path=folder + "wavenumber_full.dat",
                               spectral_resolution=case.spectral_resolution)
        grid = api.SpectralGrid(cfg, *RANGE)
        symbols = [s for s, _ in case.opacity_entries]
        folders = [f for _, f in case.opacity_entries]
        if model == "emission":
            m = api.EmissionModel(cfg, grid, LEVELS, "scm", [], symbols, folders)
        elif model == "transmission":
            m = api.TransmissionModel(cfg, grid, LEVELS, symbols, folders)
        else:
            m = api.OccultationModel(cfg, grid, LEVELS, stellar[0], stellar[1], symbols, folders)
        _OPEN[model] = (m, oracle_api.EmissionOracle(folder), folder)
    return _OPEN[model]


def _args(model, variable_gravity):
    """(a, b) of each model's calcSpectrum: emission radius and distance,
    transmission planet radius and radius ratio, eclipse radius ratio."""
    if model == "emission":
        return synth.RADIUS_JUPITER, 10 * synth.PARSEC
    if model == "transmission":
        return synth.RADIUS_JUPITER, synth.RADIUS_JUPITER / synth.RADIUS_SUN
    return 0.11, 0.0


def _engine(m, model, g, a, b, p, t, species, x, variable_gravity=False, cloud=()):
    if model == "emission":
        return m.calcSpectrum(g, a, b, p, t, species, x, cloud, variable_gravity)
    if model == "transmission":
        return m.calcSpectrum(g, a, b, p, t, species, x, cloud, variable_gravity)
    return m.calcSpectrum(g, a, p, t, species, x, cloud)


@pytest.mark.parametrize("model", MODELS)
def test_direct_grid_matches_oracle(tmp_path_factory, model):
    m, o, _ = opened(tmp_path_factory, model)
    np.testing.assert_array_equal(m.wavenumber_list, o.grid()[0])
    np.testing.assert_array_equal(m.wavelength_list, o.grid()[1])
    assert m.wavelength_list.max() <= RANGE[1] * 1.001 and m.wavelength_list.min() >= RANGE[0] * 0.999


@pytest.mark.parametrize("variable_gravity", [False, True])
@pytest.mark.parametrize("model", MODELS)
def test_direct_spectrum_matches_oracle(tmp_path_factory, model, variable_gravity):
    if model == "secondary_eclipse" and variable_gravity:
        pytest.skip("the eclipse's calcSpectrum has no variable gravity")
    m, o, _ = opened(tmp_path_factory, model)
    a, b = _args(model, variable_gravity)
    for seed in (3, 4):
        rng = np.random.default_rng(seed)
        p, t, x = synth.direct_structure(rng, LEVELS, SPECIES)
        g = 10 ** rng.uniform(2.8, 3.6)
        print(f"[gpu] direct {model} seed {seed} variable gravity {variable_gravity}", flush=True)
        spec = _engine(m, model, g, a, b, p, t, SPECIES, x, variable_gravity)
        ref = o.direct_spectrum(g, a, b, p, t, SPECIES, x, variable_gravity)
        rel = float(np.max(np.abs(spec - ref) / np.abs(ref)))
        print(f"[gpu] direct {model}: max rel err {rel:.2e}", flush=True)
        assert rel <= SPECTRUM_RTOL
        # FastChem's symbols name the same species (fixed_chemistry.cpp:42-63)
        np.testing.assert_array_equal(_engine(m, model, g, a, b, p, t, FASTCHEM, x, variable_gravity), spec)


@pytest.mark.parametrize("model", MODELS)
def test_direct_cloud_matches_oracle(tmp_path_factory, model):
    """A caller-made cloud (the reference's FixedCloudModel): layer optical
    depths added to the gas's (emission, eclipse) or, over the layer's
    thickness, to the level extinction (transmission); an all-zero cloud is
    no cloud."""
    import cases
    m, o, _ = opened(tmp_path_factory, model)
    a, b = _args(model, False)
    rng = np.random.default_rng(5)
    p, t, x = synth.direct_structure(rng, LEVELS, SPECIES)
    g = 10 ** rng.uniform(2.8, 3.6)
    cloud = cases.direct_cloud(m.wavelength_list, LEVELS)
    print(f"[gpu] direct {model} with a cloud", flush=True)
    spec = _engine(m, model, g, a, b, p, t, SPECIES, x, False, cloud)
    ref = o.direct_spectrum(g, a, b, p, t, SPECIES, x, False, cloud)
    rel = float(np.max(np.abs(spec - ref) / np.abs(ref)))
    print(f"[gpu] direct {model} cloud: max rel err {rel:.2e}", flush=True)
    assert rel <= SPECTRUM_RTOL
    clear = _engine(m, model, g, a, b, p, t, SPECIES, x, False, np.zeros_like(cloud))
    np.testing.assert_array_equal(clear, _engine(m, model, g, a, b, p, t, SPECIES, x, False))
    assert np.max(np.abs(spec / clear - 1)) > 1e-3                     # the cloud matters


def test_direct_closed_forms(tmp_path_factory):
    """Isothermal emission: pi B(T) 1e-3 (R/d)^2 at every wavenumber.  A
    pure N2 atmosphere (no absorber, no scatterer): the transit depth is the
    bare disc, (R_p/R_*)^2 1e6.  A planet isothermal at the star's
    temperature against that star (tabulated at the grid's own wavelengths,
    so interpolated exactly): (Rp/Rs)^2 1e6."""
    rng = np.random.default_rng(9)
    p, _, x = synth.direct_structure(rng, LEVELS, SPECIES)
    g = 10 ** 3.3
    m, _, _ = opened(tmp_path_factory, "emission")
    t = np.full(LEVELS, 1400.0)
    r, d = _args("emission", False)
    spec = m.calcSpectrum(g, r, d, p, t, SPECIES, x)
    nu = m.wavenumber_list
    h, c, kb = synth.PLANCK_H, synth.LIGHT_C, synth.BOLTZMANN_K
    expected = math.pi * 2 * h * c * c * nu ** 3 / np.expm1(h * c * nu / (kb * 1400.0)) * 1e-3 * r * r / d / d
    np.testing.assert_allclose(spec, expected, rtol=1e-10)

    m, _, _ = opened(tmp_path_factory, "transmission")
    rp, ratio = _args("transmission", False)
    spec = m.calcSpectrum(g, rp, ratio, p, np.full(LEVELS, 900.0), ["N2"], np.ones((LEVELS, 1)))   # no opacity
    np.testing.assert_allclose(spec, ratio ** 2 * 1e6, rtol=1e-12)

    case = synth.tiny_case()
    case.model = "secondary_eclipse"
    folder = synth.write_direct_case(case, str(tmp_path_factory.mktemp("direct_star")), RANGE)
    cfg = api.GlobalConfig(retrieval_folder_path=folder, cross_section_file_path=folder + "opacity/",
                           wavenumber_file_path=folder + "wavenumber_full.dat",
                           spectral_resolution=case.spectral_resolution)
    probe = api.EmissionModel(cfg, api.SpectralGrid(cfg, *RANGE), LEVELS, "scm", [],
                              [s for s, _ in case.opacity_entries], [f for _, f in case.opacity_entries])
    lam = probe.wavelength_list
    probe.close()
    star = api.OccultationModel(cfg, api.SpectralGrid(cfg, *RANGE), LEVELS, lam, _stellar(lam),
                                [s for s, _ in case.opacity_entries], [f for _, f in case.opacity_entries])
    spec = star.calcSpectrum(g, 0.11, p, np.full(LEVELS, T_STAR), SPECIES, x)
    np.testing.assert_allclose(spec, 0.11 ** 2 * 1e6, rtol=1e-10)
    star.close()


def test_direct_refuses_bad_input(tmp_path_factory):
    m, _, _ = opened(tmp_path_factory, "emission")
    rng = np.random.default_rng(1)
    p, t, x = synth.direct_structure(rng, LEVELS, SPECIES)
    r, d = _args("emission", False)
    with pytest.raises(api.BearError) as e:
        m.calcSpectrum(1e3, r, d, p[:-1], t[:-1], SPECIES, x[:-1])
    assert e.value.code == -1 and "levels" in str(e.value)
    with pytest.raises(api.BearError) as e:
        m.calcSpectrum(1e3, r, d, p, t, ["H2O", "Xx", "H2", "He"], x)
    assert e.value.code == -1 and "Xx" in str(e.value)
    with pytest.raises(api.BearError) as e:                                           # one depth per layer and point
        m.calcSpectrum(1e3, r, d, p, t, SPECIES, x, np.zeros((LEVELS, m.nb_spectral_points)))
    assert e.value.code == -1 and "cloud" in str(e.value)
